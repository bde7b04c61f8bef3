/*
 * Oracle C ABI — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of partapp's pictorial-structure inference path
 * (libPictStruct `--find_obj`), templated on float ("the reference CPU path")
 * and double ("the FP64 oracle").  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only as the checker or the CPU
 * baseline: the product path (libpartapp_ps.so) never links or calls it.
 *
 * Every function names the reference file:line it restates.
 */
#ifndef PS_ORACLE_H_
#define PS_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Subset of ExpParam used on the path (libPartApp/ExpParam.proto:189-195,229).
 * The proto declares these as float / uint32 and the reference mixes float and
 * double arithmetic on them, so the types are kept. */
typedef struct orc_exp_param {
  float min_part_rotation;       /* degrees, default -180 */
  float max_part_rotation;       /* degrees, default  180 */
  unsigned num_rotation_steps;   /* R */
  float min_object_scale;
  float max_object_scale;
  unsigned num_scale_steps;      /* S */
  float strip_border_detections; /* fraction of the width, < 0.5 */
} orc_exp_param;

/* object_detect::Joint after loadJoints (objectdetect.h:54-86,
 * objectdetect_aux.cpp:54-141): 0-based ids, already flipped. */
typedef struct orc_joint {
  int child_idx;
  int parent_idx;
  double offset_c[2]; /* parent_pos - child_pos, child frame  */
  double offset_p[2]; /* child_pos - parent_pos, parent frame */
  double C[4];        /* 2x2 covariance, row major            */
  double rot_mean;    /* radians */
  double rot_sigma;   /* radians */
} orc_joint;

/* status codes */
#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_TREE (-2)
#define ORC_ERR_COV (-3)
#define ORC_ERR_KERNEL (-4)
#define ORC_ERR_UNSUPPORTED (-5)

/* partapp_aux.hpp:25-55 */
double orc_value_from_index(double minval, double maxval, double num_steps, int idx);
int orc_index_from_value(double minval, double maxval, double num_steps, double val);

/* boost_math.cpp:104-117; returns the length 2k+1, or -1 if it exceeds cap */
int orc_gaussian_filter(double sigma, double *f, int cap);

/* multi_array_filter.hpp:115-155 */
int orc_filter_1d_wraparound_f32(const float *in, float *out, int n, const float *f, int f_len);
int orc_filter_1d_wraparound_f64(const double *in, double *out, int n, const double *f, int f_len);

/* multi_array_filter.hpp:211-321 (diagonal C) */
int orc_gauss_filter_diag2d_f32(const float *in, float *out, int H, int W, const double *C);
int orc_gauss_filter_diag2d_f64(const double *in, double *out, int H, int W, const double *C);

/* multi_array_transform.hpp:122-257, TM_NEAREST, fixed size; T21 is 3x3 row major */
int orc_transform_nearest_f32(const float *in, float *out, int H, int W, const double *T21, float default_value);
int orc_transform_nearest_f64(const double *in, double *out, int H, int W, const double *T21, double default_value);

/* objectdetect_findrot.cpp:292-456 — one message; child is [R][H][W] */
int orc_rot_joint_marginal_f32(const orc_exp_param *ep, const float *child, float *parent, int R, int H, int W,
                               const double *offset_c, const double *offset_p, const double *C, double rot_mean,
                               double rot_sigma, double scale, int is_sparse);
int orc_rot_joint_marginal_f64(const orc_exp_param *ep, const double *child, double *parent, int R, int H, int W,
                               const double *offset_c, const double *offset_p, const double *C, double rot_mean,
                               double rot_sigma, double scale, int is_sparse);

/* objectdetect_findrot.cpp:470-727 (+124-286).
 *   log_part_detections  [P][S][R][H][W], masked in place as the reference does
 *   log_part_posterior   [P][S][R][H][W] out (may be NULL)
 *   root_part_posterior  [S][H][W] out
 *   best_idx             [S][P][3] out: rotidx, y, x of the first flat argmax
 *   best_score           [S][P]    out */
int orc_root_posterior_rot_f32(const orc_exp_param *ep, int nparts, const int *is_detect, const int *is_upright,
                               int rootpart_idx, int njoints, const orc_joint *joints, int H, int W,
                               float *log_part_detections, float *log_part_posterior, float *root_part_posterior,
                               int *best_idx, float *best_score, int is_sparse);
int orc_root_posterior_rot_f64(const orc_exp_param *ep, int nparts, const int *is_detect, const int *is_upright,
                               int rootpart_idx, int njoints, const orc_joint *joints, int H, int W,
                               double *log_part_detections, double *log_part_posterior, double *root_part_posterior,
                               int *best_idx, double *best_score, int is_sparse);

/* objectdetect_aux.cpp:193-251: local maxima of a [D][H][W] grid, at most
 * max_n of them, each written as (d, x, y, value); returns the count. */
int orc_find_local_max_f32(const float *grid, int D, int H, int W, int max_n, double *out);

/* gaussFilter2d (multi_array_filter.hpp:374-388): diagonal C or the
 * eigenbasis path gaussFilter2dOffset (:334-369) with offset 0 */
int orc_gauss_filter2d_f32(const float *in, float *out, int H, int W, const double *C, int is_sparse);
int orc_gauss_filter2d_f64(const double *in, double *out, int H, int W, const double *C, int is_sparse);
/* boost_math.cpp:40-99; V, E row major */
int orc_eig2d(const double *C, double *V, double *E);
/* multi_array_transform.hpp:122-239: TM_DIRECT with T31 = T32 T21 (scatter),
 * TM_BILINEAR of a fixed-size output with T21 (gather through its inverse) */
int orc_transform_direct_f32(const float *in, int in_h, int in_w, float *out, int out_h, int out_w,
                             const double *T31, float default_value);
int orc_transform_bilinear_f32(const float *in, int in_h, int in_w, float *out, int out_h, int out_w,
                               const double *T21, float default_value);
/* PartApp::loadScoreGrid's grid-to-image mapping (partapp.cpp:866-899, Tig per
 * slice, NO_CLASS_VALUE 0) followed by clip_scores_fill + computeLogGrid
 * (objectdetect_findrot.cpp:834-845, objectdetect_aux.hpp:41-59) */
int orc_ingest_scores_f32(const float *score_grid, int n, int gh, int gw, const double *Tig, int interpolate, int H,
                          int W, float *out);

/* Extra unaries added to the log unaries before the sweep
 * (objectdetect_findrot.cpp:896-948), float arithmetic as the reference:
 *   setTorsoPosPrior (objectdetect_icps.cpp:137-191) on the root part's
 *     [SR][H][W] slices, params = mu_x, mu_y, var_x, var_y;
 *   addLoadDPMScore's accumulation (:469-484) on one part's [S][R][H][W] from
 *     raw scores dpm [n_rot_dpm][H][W];
 *   addDPMScore (:490-524) from log grids [n_log][H][W];
 *   computeLogGrid (multi_array_op.hpp:154-166);
 *   addExtraUnary (:526-548) over n values. */
int orc_torso_pos_prior_f32(float *grids, int SR, int H, int W, const double *params, float weight);
int orc_add_dpm_raw_f32(float *grids, int S, int R, int H, int W, const float *dpm, int n_rot_dpm, float weight);
int orc_add_dpm_log_f32(float *grids, int S, int R, int H, int W, const float *log_dpm, int n_log, float weight);
int orc_log_grid_f32(float *grid, long long n);
int orc_add_extra_unary_f32(float *grids, const float *extra, long long n, float weight);

/* The integer plan of one message, evaluated pixel by pixel with the
 * reference's own double arithmetic (objectdetect_findrot.cpp:316-359,437-448;
 * multi_array_transform.hpp:193-216).  For every rotation r the forward warp
 * reads source row floor(y - d_y + 0.5) and column floor(x - d_x + 0.5); the
 * inverse warp reads floor(y + d'_y + 0.5), floor(x + d'_x + 0.5).  The
 * returned offsets are those indices minus (y, x) at y = x = 0, and
 * *uniform is cleared when some pixel of the H x W grid disagrees.
 *   fwd_*[r] is indexed by the SOURCE slice r (its output slice is r + delta). */
int orc_message_plan(const orc_exp_param *ep, int R, int H, int W, const double *offset_c, const double *offset_p,
                     double rot_mean, double rot_sigma, double scale, int *delta, int *fwd_ox, int *fwd_oy,
                     int *inv_ox, int *inv_oy, int *uniform, int *rot_len, double *rot_taps, int *kx, int *ky);

#ifdef __cplusplus
}
#endif
#endif
