/*
 * partapp_ps — MI355X (gfx950) pictorial-structure inference, C ABI.
 *
 * Drop-in replacement for the rotation-aware sum-product sweep of partapp's
 * libPictStruct (the `--find_obj` stage).  Every entry point takes plain
 * pointers and sizes; device buffers belong to the caller, plans and their
 * scratch belong to the library.  All functions return 0 on success or a
 * negative ps_status; ps_last_error_string() describes the last failure of
 * the calling thread.  Nothing aborts: the asserts of the reference become
 * status codes.
 *
 * Layouts (C order, x fastest, as boost::multi_array in
 * libMultiArray/multi_array_def.h:36 and objectdetect_findrot.cpp:494-495):
 *   log unaries / log marginals of one instance: [P][S][R][H][W] float
 *   root posterior of one instance:              [S][H][W]       float
 *   MAP records:                                 [inst][S][P]    ps_map
 */
#ifndef PARTAPP_PS_H_
#define PARTAPP_PS_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ps_status {
  PS_OK = 0,
  PS_ERR_INVALID_ARG = -1,   /* null pointer, bad size, bad flag            */
  PS_ERR_TREE = -2,          /* not one root + chains (objectdetect_findrot.cpp:210,786;
                                objectdetect_aux.cpp:131)                      */
  PS_ERR_COVARIANCE = -3,    /* C diagonal not > 0 (multi_array_filter.hpp:219) */
  PS_ERR_KERNEL_SIZE = -4,   /* Gaussian kernel >= 1000 taps (multi_array_filter.hpp:243-244) */
  PS_ERR_UNSUPPORTED = -5,   /* e.g. R > 64, more than 16 root children      */
  PS_ERR_WARP = -6,          /* a nearest-neighbour warp's offset changes more than once along an axis */
  PS_ERR_HIP = -7,           /* HIP runtime failure                          */
  PS_ERR_NO_DEVICE = -8,     /* no usable gfx950 device                      */
  PS_ERR_OUT_OF_MEMORY = -9,
  PS_ERR_IO = -10,           /* file cannot be read / written, or is not a level-5 MAT file */
  PS_ERR_NOT_FOUND = -11     /* no such MAT-file variable                    */
} ps_status;

/* ExpParam fields used on the path (libPartApp/ExpParam.proto:189-195,229);
 * float / uint32 as in the proto because the reference's arithmetic depends on it. */
typedef struct ps_exp_param {
  float min_part_rotation;       /* degrees (default -180)          */
  float max_part_rotation;       /* degrees (default  180)          */
  unsigned num_rotation_steps;   /* R (default 48), 1..64           */
  float min_object_scale;
  float max_object_scale;
  unsigned num_scale_steps;      /* S                               */
  float strip_border_detections; /* fraction of width, 0 <= f < 0.5 */
} ps_exp_param;

/* object_detect::Joint after loadJoints (objectdetect.h:54-86,
 * objectdetect_aux.cpp:54-141): ids 0-based, flip already applied. */
typedef struct ps_joint {
  int child_idx;
  int parent_idx;
  double offset_c[2]; /* parent_pos - child_pos in the child frame (pixels) */
  double offset_p[2]; /* child_pos - parent_pos in the parent frame         */
  double C[4];        /* 2x2 covariance, row major                          */
  double rot_mean;    /* radians */
  double rot_sigma;   /* radians */
} ps_joint;

/* PartConfig + ExpParam subset of one model (PartConfig.proto:1-59). */
typedef struct ps_model_desc {
  int num_parts;          /* P */
  const int *is_detect;   /* [P] PartDef.is_detect (default 1)  */
  const int *is_upright;  /* [P] PartDef.is_upright (default 0) */
  const int *is_root;     /* [P] PartDef.is_root; the root is the one part with
                             is_detect && is_root (objectdetect_findrot.cpp:781-786) */
  int num_joints;         /* must be P - 1 */
  const ps_joint *joints; /* default spatial model, used by instances with joints == NULL */
  ps_exp_param exp_param;
  int is_sparse;          /* bIsSparse of the upward pass; only the off-diagonal
                             covariance path reads it */
} ps_model_desc;

typedef struct ps_model ps_model; /* opaque, immutable after creation        */
typedef struct ps_plan ps_plan;   /* opaque: tables + scratch for one batch   */

/* One image of a batch.  Offsets are in elements of the arrays passed to
 * ps_plan_execute / ps_marginals, so instances of uneven size can be packed. */
typedef struct ps_instance {
  int height;                 /* H */
  int width;                  /* W */
  long long unary_offset;     /* start of this instance's [P][S][R][H][W] log unaries   */
  long long marginal_offset;  /* start of its [P][S][R][H][W] log marginals (output)    */
  long long root_offset;      /* start of its [S][H][W] root posterior (output)         */
  const ps_joint *joints;     /* NULL: the model's joints; else num_joints joints with the
                                 model's topology (image-conditioned spatial model)      */
} ps_instance;

/* MAP of one (instance, scale, part): the first flat argmax of the log
 * marginal, objectdetect_findrot.cpp:255-277 (PartHyp, objectdetect.h:88-195). */
typedef struct ps_map {
  int rotidx;
  int y;
  int x;
  float score;
} ps_map;

/* ------------------------------------------------------------------ models */
int ps_model_create(const ps_model_desc *desc, ps_model **out);
void ps_model_destroy(ps_model *model);

/* ------------------------------------------------------------------- plans
 * A plan fixes the batch (sizes, per-instance joints) and owns the device
 * tables and the scratch of the sweep; executing it allocates nothing.
 * Replaces the per-scale, per-message allocations of computeRootPosteriorRot
 * (objectdetect_findrot.cpp:560-578, 339-340, 373). */
int ps_plan_create(const ps_model *model, const ps_instance *instances, int n_instances, ps_plan **out);
void ps_plan_destroy(ps_plan *plan);
size_t ps_plan_scratch_bytes(const ps_plan *plan);
/* Launch the whole sweep on `stream` (a hipStream_t, NULL = default stream).
 * log_unary is read-only: the upright / border masks of
 * objectdetect_findrot.cpp:509-551 are applied on the fly instead of being
 * written into the caller's grids.  log_marginal also holds intermediate
 * grids while the sweep runs (a part's marginal grid carries the root's
 * message to it until that part's marginal is written); it holds the
 * marginals once the sweep's work on the stream is done.  map_dev:
 * [n_instances][S][P] ps_map in device memory.  Asynchronous.
 * root_posterior may be NULL. */
int ps_plan_execute(ps_plan *plan, const float *log_unary, float *log_marginal, float *root_posterior,
                    ps_map *map_dev, void *stream);
/* Same, through a HIP graph captured on first use for this pointer set
 * (re-captured when the pointers or the stream change). */
int ps_plan_execute_graph(ps_plan *plan, const float *log_unary, float *log_marginal, float *root_posterior,
                          ps_map *map_dev, void *stream);

/* Arithmetic of the spatial Gaussian passes (the rest of every message stays
 * FP32).  PS_METHOD_FP32 (default) is the drop-in: sequential FP32 dot
 * products in the reference's order.  The other methods run the x, rotation
 * and y filters as banded-Toeplitz / circulant GEMMs on the matrix cores with
 * FP32 accumulation:
 *   x filter       BF16: data and taps rounded to bf16 (one product);
 *                  BF16X2: data split into bf16 hi + lo (two products);
 *                  BF16X3: data and taps split (three products)
 *   rotation       data and taps split into bf16 hi + lo (three products)
 *   intermediate   BF16 / BF16X2: fp16, scaled by a power of two chosen from
 *                  the filters' absolute sums; BF16X3: FP32
 *   y filter       BF16 / BF16X2: fp16 data (exact), taps split into fp16
 *                  hi + lo (two products); BF16X3: data and taps split into
 *                  bf16 hi + lo with a power-of-two rescale per column of a
 *                  rotation slice (three products)
 * PS_METHOD_BF16_MIXED (the benchmarked choice) runs the kernels of
 * PS_METHOD_BF16.  Relative L1 error of the normalised marginals against FP64
 * on 300x300x24 grids: about 1.5e-4 for BF16 / BF16_MIXED, 1e-4 for BF16X2,
 * 2e-6 for BF16X3 and FP32.  Available for 24 rotations and half widths up to
 * 47 (sigma * scale up to 15.5 px); a level of messages outside that runs FP32
 * (ps_plan_method_levels counts the levels that use the method).  No
 * reference counterpart: a precision option of this library. */
enum {
  PS_METHOD_FP32 = 0,
  PS_METHOD_BF16 = 1,
  PS_METHOD_BF16X2 = 2,
  PS_METHOD_BF16X3 = 3,
  PS_METHOD_BF16_MIXED = 4
};
int ps_plan_set_method(ps_plan *plan, int method);
/* Launch-shape knobs for an autotuner (results do not depend on them):
 *   PS_TUNE_MFMA_ROWS        output rows per block of the BF16 x pass (1..64,
 *                            default 4);
 *   PS_TUNE_FP32_ROW_THREADS threads per block of the FP32 row pass (a
 *                            multiple of 64 up to 512; 0 = one per column,
 *                            the default);
 *   PS_TUNE_FP32_COL_ROWS    output rows per lane of the FP32 column pass
 *                            (8 or 16; 0 = 8, the default).
 * ps_plan_tune sets a knob for every launch class; ps_plan_tune_class for
 * one class.  A class is one (level, width class) row-pass launch of the
 * sweep, in execution order: ps_plan_classes returns their count and, per
 * class, (level, jobs, tallest grid, widest grid) in info[4 * c ..]; the
 * column-pass knob of a class reaches the column-pass launches of its level
 * and width class.  No reference counterpart. */
enum { PS_TUNE_MFMA_ROWS = 1, PS_TUNE_FP32_ROW_THREADS = 2, PS_TUNE_FP32_COL_ROWS = 3 };
int ps_plan_tune(ps_plan *plan, int knob, int value);
int ps_plan_tune_class(ps_plan *plan, int knob, int cls, int value);
int ps_plan_classes(const ps_plan *plan, int *n_classes, int *info, int max_classes);
int ps_plan_method_levels(const ps_plan *plan, int *levels, int *mfma_levels);

/* MAP guard of the BF16 methods (no counterpart in the reference, whose MAP
 * is the FP32 argmax).  With tau > 0 and a BF16 method, every execution also
 * keeps, per (instance, scale, part), the best and second best log-marginal
 * values; an instance where some best does not beat its second best by more
 * than tau (twice the method's bound on the error of a log-marginal value)
 * is swept again with the FP32 passes, on the device, before the execution
 * ends (its marginals, root posterior and MAP records then come from FP32).
 * With ps_plan_execute the flags are read back after the certificate (the
 * call then waits for the stream) and only the flagged instances' jobs are
 * launched again; under ps_plan_execute_graph the flags stay on the device
 * and the FP32 launches cover the batch, their other blocks exiting at once.
 * tau = 0 turns the guard off (the default).  ps_plan_read_map_guard
 * reports how many instances the last execution re-evaluated (and, if
 * flags_host is given, which: [n_instances]); it waits for `stream`. */
/* Pruned execution (no counterpart in the reference).  keep_dev holds one
 * byte per 64 x 64 tile of every (instance, part, scale, rotation) plane:
 * instance i's flags start at byte instance_offsets[i] of
 * ps_plan_keep_layout and are laid out [P][S][R][ceil(H/64)][ceil(W/64)].
 * In the BF16 methods' column pass a destination tile whose flag is 0 is
 * not computed and is written as LOG_ZERO: the sweep then runs on the
 * pruned state space (zero probability outside the kept tiles), whose
 * marginals differ from the exact ones by at most twice the probability mass
 * the exact marginals put on the dropped tiles.  Levels that run the FP32
 * passes compute every tile.  fallback_dev (optional, [n_instances]): an
 * instance with a nonzero flag runs the whole (unpruned) sweep instead.
 * Asynchronous; the MAP guard is not applied. */
int ps_plan_keep_layout(const ps_plan *plan, long long *instance_offsets, long long *total_bytes);
int ps_plan_execute_pruned(ps_plan *plan, const float *log_unary, float *log_marginal, float *root_posterior,
                           ps_map *map_dev, const unsigned char *keep_dev, const unsigned *fallback_dev,
                           void *stream);

int ps_plan_set_map_guard(ps_plan *plan, float tau);
int ps_plan_read_map_guard(ps_plan *plan, int *n_reevaluated, unsigned *flags_host, void *stream);

/* Per-kernel timing of a plan's executions, for measurement: while enabled,
 * every kernel launch of ps_plan_execute (a level's eigenbasis stages count as
 * one) is bracketed by HIP events on its stream.
 * ps_plan_read_profile waits for the recorded launches and returns, per kernel
 * class (0 reset, 1 leaf max, 2 row pass, 3 column pass, 4 root combine,
 * 5 MAP decode), the summed milliseconds and the launch count since the last
 * read; the counters then restart.  Not available with ps_plan_execute_graph. */
int ps_plan_set_profiling(ps_plan *plan, int enable);
/* Same, bracketing only the launches of the classes in `mask` (bit c = class
 * c, 0 = off): fewer events on the stream while one kernel is measured. */
int ps_plan_set_profiling_classes(ps_plan *plan, unsigned mask);
int ps_plan_read_profile(ps_plan *plan, double ms[6], int launches[6]);

/* ----------------------------------------------------- reference-shaped API
 * computeRootPosteriorRot (objectdetect_findrot.cpp:470-727; objectdetect.h:269-274)
 * for a batch: plan, execute, copy the MAP records to host memory
 * (map_host: [n_instances][S][P]) and wait.  Device pointers as above. */
int ps_marginals(const ps_model *model, const ps_instance *instances, int n_instances, const float *log_unary,
                 float *log_marginal, float *root_posterior, ps_map *map_host, void *stream);

/* computeRotJointMarginal (objectdetect_findrot.cpp:292-456; objectdetect.h:277-282):
 * one message child[R][H][W] -> parent[R][H][W] on the device, with the
 * model's exp_param.  is_sparse is bIsSparse: with an off-diagonal C it
 * selects the nearest scatter (else bilinear) rotation into the eigenbasis
 * (multi_array_filter.hpp:351-355).  Synchronous. */
int ps_message(const ps_model *model, const float *child_dev, float *parent_dev, int height, int width,
               const double offset_c[2], const double offset_p[2], const double C[4], double rot_mean,
               double rot_sigma, double scale, int is_sparse, void *stream);

/* ps_message without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * message's tables, keys and scratch grids live in a device array of
 * PS_MESSAGE_NOALLOC_BYTES that the library holds on every GPU it runs on
 * until the process ends; a message that needs more (about
 * R * height * ceil32(width) * 4 bytes plus a few KB of tables, plus
 * 12 * R * th * tw bytes of eigenbasis grids for an off-diagonal C) returns
 * PS_ERR_UNSUPPORTED.  Same arguments, results and errors as ps_message;
 * synchronous.  Calls on one GPU share the array: the library serialises
 * them with a process-wide lock (overlapping calls would race on it). */
#define PS_MESSAGE_NOALLOC_BYTES (32u << 20)
int ps_message_noalloc(const ps_model *model, const float *child_dev, float *parent_dev, int height, int width,
                       const double offset_c[2], const double offset_p[2], const double C[4], double rot_mean,
                       double rot_sigma, double scale, int is_sparse, void *stream);

/* findLocalMax (objectdetect_aux.cpp:193-251) on device grids.  Grid g is
 * [D][H][W] floats at grids_dev + offsets[g] (dims[3g..3g+2] = D, H, W; D is
 * rotations for a part marginal, scales for the root posterior).  A point is
 * kept when no spatial 8-neighbour in its slice is strictly greater and the
 * neighbouring slices are strictly lower there.  At most max_n points per
 * grid: in scan order (d, then x, then y) when there are at most max_n,
 * otherwise the max_n largest by value, in decreasing order (ties in scan
 * order).  out_dev: [n_grids][max_n] records, count_dev: [n_grids].
 * max_n <= 4096.  Allocates a workspace of 8 bytes per grid point; synchronous.
 * Feeds best_part_hyp (objectdetect_findrot.cpp:280-283) and the root
 * HypothesisList (:1037-1048). */
typedef struct ps_local_max {
  int d;  /* slice: rotation index (part marginal) or scale index (root posterior) */
  int x;
  int y;
  float score;
} ps_local_max;
int ps_find_local_max(const float *grids_dev, const long long *offsets, const int *dims, int n_grids, int max_n,
                      ps_local_max *out_dev, int *count_dev, void *stream);

/* Unary ingestion (PartApp::loadScoreGrid, partapp.cpp:866-899, followed by
 * clip_scores_fill + computeLogGrid, objectdetect_findrot.cpp:834-845):
 * n score-grid slices [n][grid_height][grid_width] on the device, each mapped
 * to image coordinates with its homogeneous 3x3 Tig (host, [n][9], row
 * major: image = Tig * grid) by nearest scatter (interpolate = 0, TM_DIRECT)
 * or bilinear gather (TM_BILINEAR), with NO_CLASS_VALUE 0 where nothing
 * lands; then scores < 0 become 1e-4 and the log is taken (0 -> LOG_ZERO).
 * Writes log_unary_dev [n][height][width].  Allocates a workspace (8 bytes
 * per output cell for the scatter); synchronous. */
int ps_ingest_scores(const float *score_grid_dev, int n, int grid_height, int grid_width, const double *Tig,
                     int interpolate, int height, int width, float *log_unary_dev, void *stream);

/* Extra unaries: the additions findObjectImageRotJoints makes to the log
 * unaries after ingestion and before the sweep (objectdetect_findrot.cpp:
 * 896-948), in place on device grids, with the reference's arithmetic (float
 * values; exp of the position prior in double; additions of a double term
 * done in double and rounded once to float).  Asynchronous on `stream`; no
 * allocation.
 *
 * setTorsoPosPrior (objectdetect_icps.cpp:137-191): root_unary_dev holds the
 * root part's n_slices = S * R grids [n_slices][height][width]; params =
 * (mu_x, mu_y, var_x, var_y) of torso_pos_prior.mat (getTorsoPosPriorParams,
 * :32-45); adds log(max(1e-4, weight * exp(-(cx - x - mu_x)^2 / (2 var_x)) *
 * exp(-(cy - y - mu_y)^2 / (2 var_y)))) with (cx, cy) = (width / 2, height / 2). */
int ps_add_torso_pos_prior(float *root_unary_dev, int n_slices, int height, int width, const double params[4],
                           float weight, void *stream);

/* DPM scores for one part (part_unary_dev: [S][R][height][width]) from
 * dpm_dev [n_dpm][height][width]:
 *   PS_DPM_RAW        addLoadDPMScore's accumulation (objectdetect_icps.cpp:469-484)
 *                     on raw scores v: += (v > 1e-4 ? weight * log(v) : log(1e-4));
 *                     slice r of dpm when n_dpm == R, else slice 0;
 *   PS_DPM_LOG        addDPMScore (:490-524) on log grids: += weight * g;
 *                     slice r when n_dpm > 1 (then n_dpm >= R), else slice 0;
 *                     weight must be > 0 (:501);
 *   PS_DPM_LOG_OF_RAW computeLogGrid of raw scores (0 -> LOG_ZERO;
 *                     multi_array_op.hpp:154-166) then PS_DPM_LOG, the DPM
 *                     torso path (objectdetect_findrot.cpp:896-903).
 * Loading the scoregrid MAT files (loadDPMScoreGrid, :550-575) is the
 * caller's, e.g. through ps_mat_var. */
enum { PS_DPM_RAW = 0, PS_DPM_LOG = 1, PS_DPM_LOG_OF_RAW = 2 };
int ps_add_dpm_score(float *part_unary_dev, int S, int R, int height, int width, const float *dpm_dev, int n_dpm,
                     float weight, int mode, void *stream);

/* addExtraUnary (objectdetect_icps.cpp:526-548): log_unary_dev[i] +=
 * weight * extra_dev[i] for n values (all parts, scales and rotations). */
int ps_add_extra_unary(float *log_unary_dev, const float *extra_dev, long long n, float weight, void *stream);

/* --------------------------------------------------------- MAT-file (level 5)
 * The on-disk formats around the path: joint_<child>_<parent>.mat
 * (objectdetect_learnparam.cpp:61-179), the cell_scoregrid / transform_Ti2 /
 * transform_T2g score-grid files (partapp.cpp:792-857) and the log_prob_grid /
 * best_conf dumps (objectdetect_findrot.cpp:239-253, 1005-1011), read and written
 * like matlab_io (matlab_io.hpp:64-230, matlab_cell_io.hpp:27-130): numeric
 * arrays of any storage type are returned as double in C order; cell arrays
 * are addressed by their column-major cell index.  Host code, no device. */
typedef struct ps_mat ps_mat;
int ps_mat_open(const char *path, ps_mat **out);
void ps_mat_close(ps_mat *mat);
/* Variable `name` (or element cell_index >= 0 of that cell array): dims in
 * MATLAB order, and, when out is not NULL, numel values in C order. */
int ps_mat_var(const ps_mat *mat, const char *name, long long cell_index, int *is_cell, int *ndim,
               long long dims[8], double *out, long long capacity);
/* Write (append != 0: add to an existing file) one real array given in C
 * order as class double or single, optionally zlib-compressed (matOpen "wz"). */
int ps_mat_write(const char *path, int append, const char *name, const double *data, int ndim,
                 const long long *dims, int single, int compress);
/* load_joint (objectdetect_learnparam.cpp:126-170) + the per-joint part of
 * loadJoints (objectdetect_aux.cpp:96-136): flipped if asked, 0-based ids. */
enum { PS_JOINT_POS_GAUSSIAN = 1, PS_JOINT_ROT_GAUSSIAN = 2 };
int ps_load_joint_mat(const char *path, int flip, ps_joint *out, int *type);

/* --------------------------------------------------------------- utilities */
const char *ps_last_error_string(void);
const char *ps_version(void);

/* The integer plan of one message as the library computes it on the host:
 * rotation shift delta, per-rotation warp offsets (forward indexed by source
 * slice), the rotation taps (rot_len 0 when rot_sigma <= 0) and the half
 * widths of the spatial filters for covariance C.  For tests. */
int ps_debug_message_plan(const ps_exp_param *ep, int height, int width, const double offset_c[2],
                          const double offset_p[2], const double C[4], double rot_mean, double rot_sigma,
                          double scale, int *delta,
                          int *fwd_ox, int *fwd_oy, int *inv_ox, int *inv_oy, int *rot_len, float *rot_taps,
                          int *kx, int *ky);

/* Seeded synthetic instance (the benchmark workload): log unaries
 * N(-3, 0.5) plus 3 Gaussian bumps (amplitude 4, std 6 px) per part at seeded
 * (x, y, rotation bin), and joints with offsets U(-40,40), sigma_x,y U[4,12],
 * rot_sigma U[0.1,0.4], rot_mean U[-pi/6,pi/6] (0 when zero_rot_mean).
 * tree: 0 = the 10-part person tree (root = part 4, torso).
 * Host memory; unary: [P][S][R][H][W], joints: [P-1].  Deterministic in seed. */
int ps_synth_joints(unsigned long long seed, int tree, int zero_rot_mean, ps_joint *joints, int *num_joints);
int ps_synth_unary(unsigned long long seed, int num_parts, int num_scales, int num_rot, int height, int width,
                   float *unary, int n_threads);
/* The same generator on the device (tens of GB of batches): writes the
 * [P][S][R][H][W] grid at unary_dev; synchronous on `stream`.  Values equal
 * the host generator's up to the device's double-precision libm. */
int ps_synth_unary_dev(unsigned long long seed, int P, int S, int R, int H, int W, float *unary_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
