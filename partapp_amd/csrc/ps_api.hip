// C ABI of the gfx950 pictorial-structure sweep (include/partapp_ps.h).
//
// A plan turns a batch of instances into device tables: one MsgJob per
// (instance, scale, joint, direction), one RootJob per (instance, scale),
// MaxJobs for the leaves and MapJobs for the MAP records.  Executing the plan
// launches, in stream order,
//   reset keys -> leaf maxima -> [rowpass, colpass] per upward level
//   -> root combine -> [rowpass, colpass] per downward level -> MAP decode,
// every launch covering all instances, scales and chains of its level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ps_internal.hpp"
#include "ps_message.hpp"
#include "ps_plan.hpp"

namespace ps {
extern thread_local std::string g_last_error;
hipError_t launch_reset(const Bases &b, int n_keys, int n_map, int n_inst, unsigned *flags, int n_flags,
                        hipStream_t s);
hipError_t launch_invert_flags(const unsigned *in, unsigned *out, int n, hipStream_t s);
hipError_t launch_certify(const unsigned long long *top2, int n_map, int n_inst, float tau, unsigned *flags,
                          hipStream_t s);
hipError_t launch_max(const Bases &b, const MaxJob *jobs, int n_jobs, int R, int max_hw, hipStream_t s);
size_t rowpass_lds_bytes(int R, int W, int kx);
hipError_t rowpass_set_lds_limit(size_t bytes);
hipError_t colpass_set_lds_limit(size_t bytes);
hipError_t launch_rowpass(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                          int R, int max_h, int max_w, size_t lds_bytes, hipStream_t s, int threads, bool split);
size_t colpass_lds_bytes(int ky, int rows);
hipError_t launch_colpass(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                          int R, int epi_hi, int max_w, size_t lds_bytes_16, hipStream_t s, int rows, bool split);
hipError_t launch_offdiag(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts, int R,
                          long long max_cells, long long max_hw, hipStream_t s);
hipError_t launch_root(const Bases &b, const RootJob *jobs, int n_jobs, int k, int R, int max_hw, bool lse64,
                       hipStream_t s);
size_t rowpass_mfma_lds_bytes(int R, int W, int kx, int KK, int terms);
size_t colpass_mfma_lds_bytes(int KK, int terms);
size_t colpass_f16_lds_bytes(int KK, int H);
hipError_t mfma_set_lds_limit(size_t row_bytes, size_t col_bytes);
hipError_t launch_rowpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                               const uint4 *frags, int R, int terms, int rows, int max_h, int max_w,
                               size_t lds_bytes, bool leaf, hipStream_t s);
hipError_t launch_colpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const uint4 *frags, const int *shifts,
                               int R, int terms, bool t16, int epi_hi, int max_w, size_t lds_bytes, bool leaf,
                               hipStream_t s);
hipError_t launch_map_decode(const Bases &b, const MapJob *jobs, int n, ps_map *out, hipStream_t s);
}  // namespace ps

using namespace ps;

struct ps_model {
  Model m;
};

namespace {

// One level of messages: jobs [off, off + count), the first n_diag with a
// diagonal covariance (row pass + column pass), the rest off-diagonal (row
// pass + eigenbasis stages).
// Diagonal jobs of a level with the same count of 64-column groups
// (ceil(W / 64)), contiguous after sorting: each class gets its own row-pass
// and column-pass launches, sized for its widest job, so that the blocks of
// narrow jobs in an uneven batch do not carry idle waves.
struct WidthClass {
  int start = 0, count = 0, max_h = 0, max_w = 0;
  int epi_hi = kEpiUpRoot;  // highest epilogue kind of its jobs
  int mfma_rows = 0;        // tuned rows per block of the BF16 x pass (0: the plan's)
  int row_threads = 0;      // tuned threads per block of the FP32 row pass (0: one per column)
  int col_rows = 0;         // tuned output rows per lane of the FP32 column pass (0: the default, 8)
  bool leaf = false;        // a leaf's message among its jobs (its source is a unary)
  bool split = false;       // a job whose warp changes its offset along an axis (MsgJob::split)
  // the jobs of each instance: consecutive within the class (the level's
  // jobs keep instance order within a class); (instance, first job relative
  // to the level, count), by instance
  std::vector<int> run_inst, run_first, run_count;
};

// record the instance runs of a class of a level's jobs v
void class_runs(WidthClass &c, const std::vector<MsgJob> &v) {
  for (int j = c.start; j < c.start + c.count; ++j) {
    if (v[j].src_kind == kSrcUnary) c.leaf = true;
    if (v[j].split) c.split = true;
    if (c.run_inst.empty() || c.run_inst.back() != v[j].inst) {
      c.run_inst.push_back(v[j].inst);
      c.run_first.push_back(j);
      c.run_count.push_back(0);
    }
    ++c.run_count.back();
  }
}

struct Level {
  int off = 0, count = 0, n_diag = 0, max_h = 0, max_w = 0;
  std::vector<WidthClass> classes;      // row passes: over the diagonal jobs [0, n_diag)
  std::vector<WidthClass> col_classes;  // column passes: the same split further by epilogue kind
  size_t lds = 0;      // row pass
  size_t col_lds = 0;  // column pass
  long long od_cells = 0, od_hw = 0;  // max R*th*tw and H*W of the off-diagonal jobs
  bool mfma_ok = false;                // every diagonal job fits the BF16 Toeplitz passes
  size_t mfma_row_lds[3] = {0, 0, 0};  // 1, 2 and 3 terms
  size_t mfma_col_lds[3] = {0, 0, 0};
  size_t mfma_col16_lds = 0;           // column pass on an fp16 row-pass output
  WidthClass od;                       // the off-diagonal jobs [n_diag, count), for their instance runs
  bool split = false;                  // some job's warp changes its offset along an axis
};

// Host bf16 of a float, round to nearest even (finite inputs).
uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_float(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// k-steps of 16 of a 32-wide Toeplitz tile with half width k
int mfma_kk(int k) { return (32 + 2 * k + 15) / 16; }

int hip_error(hipError_t e, const char *what) {
  return set_error(PS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr size_t kLdsLimit = 160 * 1024;  // LDS per workgroup on gfx950

int check_device(int &device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(PS_ERR_NO_DEVICE, "no HIP device");
  if (hipGetDevice(&device) != hipSuccess) return set_error(PS_ERR_NO_DEVICE, "hipGetDevice failed");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return set_error(PS_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(PS_ERR_NO_DEVICE, std::string("library is built for gfx950, device is ") + prop.gcnArchName);
  return PS_OK;
}

GridRef gref(int buf, long long off, int pitch) {
  GridRef g;
  g.buf = buf;
  g.pitch = pitch;
  g.off = off;
  return g;
}
const GridRef kNone = {-1, 0, 0};

}  // namespace

struct ps_plan {
  Model model;
  int device = 0;
  int n_inst = 0, S = 0, P = 0, R = 0, items = 0;
  // device tables
  void *tables = nullptr;
  MsgJob *jobs = nullptr;
  RootJob *roots = nullptr;
  MaxJob *maxes = nullptr;
  MapJob *maps = nullptr;
  float *taps = nullptr;
  int *shifts = nullptr;
  uint4 *frags = nullptr;
  int terms = 0;      // x pass: 0 FP32, else the BF16 Toeplitz terms (1, 2, 3)
  int col_terms = 0;  // y pass
  int mfma_rows = 4;  // output rows per block of the BF16 x pass (tunable)
  float *scratch = nullptr;
  size_t scratch_elems = 0;
  unsigned *keys = nullptr;
  int n_keys = 0;
  unsigned long long *map_keys = nullptr;
  int n_map = 0;
  // MAP guard of the BF16 methods: top-2 value keys per argmax key, a flag
  // per instance, and the certificate threshold (0: off)
  unsigned long long *top2 = nullptr;
  unsigned *flags = nullptr;
  float guard_tau = 0.0f;
  // pruned execution: keep flags per 64 x 64 tile, [inst][P][S][R][ty][tx]
  std::vector<long long> keep_offsets;  // per instance, bytes
  long long keep_bytes = 0;
  unsigned *not_fallback = nullptr;     // per instance: 1 - fallback flag (the pruned pass's active set)
  unsigned *flags_host = nullptr;       // pinned copy of the guard flags (eager execution)
  int max_children = 0;                 // root children (MaxJobs per item)
  // every level runs on the BF16 passes (no off-diagonal job): with an fp16
  // row-pass output no kernel reads the leaves' maxima (the row pass takes a
  // per-row exponent instead), so those sweeps skip k_max
  bool all_mfma = false;
  // the flagged instances' jobs gathered for the guard's FP32 re-sweep:
  // host copies of the tables, a pinned staging block and its device twin
  std::vector<MsgJob> h_jobs;
  std::vector<RootJob> h_roots;
  std::vector<MaxJob> h_maxes;
  std::vector<MapJob> h_maps;
  char *cstage = nullptr;
  char *ctables = nullptr;
  size_t cbytes = 0;
  std::vector<Level> up, down;
  int n_roots = 0, n_maxes = 0;
  int root_children = 0;  // the same for every root job
  int max_rows = 0, max_hw = 0;
  // graph cache
  hipGraphExec_t gexec = nullptr;
  const void *gptr[4] = {nullptr, nullptr, nullptr, nullptr};
  hipStream_t gstream = nullptr;
  // profiling: event pairs per launch of the classes in the mask, pooled
  unsigned profile_mask = 0u;
  std::vector<hipEvent_t> events;
  std::vector<int> event_class;  // class of each recorded pair
  size_t n_recorded = 0;         // pairs recorded since the last read
};

namespace {
constexpr int kProfileClasses = 6;

// Record the start event of a launch of class `cls`; returns the pair index or -1.
long profile_begin(ps_plan *pl, int cls, hipStream_t s) {
  if (!pl || !((pl->profile_mask >> cls) & 1u)) return -1;
  size_t i = pl->n_recorded;
  while (pl->events.size() < 2 * (i + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    pl->events.push_back(e);
  }
  if (pl->event_class.size() < i + 1) pl->event_class.resize(i + 1);
  pl->event_class[i] = cls;
  if (hipEventRecord(pl->events[2 * i], s) != hipSuccess) return -1;
  pl->n_recorded = i + 1;
  return (long)i;
}
void profile_end(ps_plan *pl, long i, hipStream_t s) {
  if (i >= 0) (void)hipEventRecord(pl->events[2 * i + 1], s);
}
}  // namespace

// ------------------------------------------------------------------ models

extern "C" int ps_model_create(const ps_model_desc *desc, ps_model **out) {
  if (!out) return set_error(PS_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  ps_model *m = new ps_model();
  int st = build_model(desc, m->m);
  if (st != PS_OK) {
    delete m;
    return st;
  }
  *out = m;
  return PS_OK;
}

extern "C" void ps_model_destroy(ps_model *model) { delete model; }

// ------------------------------------------------------------------- plans

namespace {

struct Builder {
  const Model &m;
  std::vector<MsgJob> jobs;
  std::vector<RootJob> roots;
  std::vector<MaxJob> maxes;
  std::vector<MapJob> maps;
  std::vector<float> taps;
  std::vector<int> shifts;
  std::vector<uint16_t> frags;  // bf16 Toeplitz fragments, 8 per 16-byte unit
  explicit Builder(const Model &mm) : m(mm) {}

  // Fragment tables of one filter for the BF16 passes: lane l at step kk holds
  // f[16 kk + 8 (l / 32) + t - l % 32], t = 0..7 (zero off the band); the
  // bf16 hi parts, then the lo parts (f - hi).  Returns the offset in 16-byte units.
  int add_frags(const std::vector<float> &f, int KK) {
    const int off = (int)(frags.size() / 8);
    const int L = (int)f.size();
    std::vector<uint16_t> lo;
    lo.reserve((size_t)KK * 64 * 8);
    for (int kk = 0; kk < KK; ++kk)
      for (int l = 0; l < 64; ++l)
        for (int t = 0; t < 8; ++t) {
          const int i = 16 * kk + 8 * (l / 32) + t - l % 32;
          const float v = (i >= 0 && i < L) ? f[i] : 0.0f;
          const uint16_t hi = bf16_rne(v);
          frags.push_back(hi);
          lo.push_back(bf16_rne(v - bf16_float(hi)));
        }
    frags.insert(frags.end(), lo.begin(), lo.end());
    return off;
  }

  // The y taps in fp16 (hi, then lo = f - hi) for the column pass that runs
  // on the fp16 MFMA, as the 8 shifted copies of f16_tap_copy_len: part p,
  // copy j, element e holds tap j + e - 32 (zero outside the filter).
  int add_taps16(const std::vector<float> &f, int KK) {
    const int off = (int)(frags.size() / 8);
    const int L = (int)f.size(), Q = f16_tap_copy_len(KK);
    std::vector<uint16_t> lo;
    for (int j = 0; j < 8; ++j)
      for (int e = 0; e < Q; ++e) {
        const int i = j + e - 32;
        const float v = (i >= 0 && i < L) ? f[i] : 0.0f;
        const _Float16 hi = (_Float16)v;
        const _Float16 lv = (_Float16)(v - (float)hi);
        frags.push_back(__builtin_bit_cast(uint16_t, hi));
        lo.push_back(__builtin_bit_cast(uint16_t, lv));
      }
    frags.insert(frags.end(), lo.begin(), lo.end());
    return off;
  }

  // Rotation filter of a message as the A operand of the row pass's second
  // GEMM, whose B operand is the x-filter accumulator itself: B position
  // p = 16 kk + 8 h + t (lane half h, element t) holds slice
  // 16 kk + 8 (t / 4) + 4 h + t % 4, so lane l, step kk, element t of the
  // table is F[l % 32][that slice] with F[ro][s] the weight of source slice s
  // in output slice ro (circular taps, objectdetect_findrot.cpp:377-420);
  // zero for slices >= R.  The hi block, then the lo block.  With 48 slices
  // there are two output tiles (ro = 32 mt + l % 32) and three k-steps: step
  // ks reads the accumulator of slice tile ks / 2, so its B position holds
  // slice 32 (ks / 2) + 16 (ks % 2) + 8 (t / 4) + 4 h + t % 4; the table is
  // [mt][ks] per block.
  int add_rot_frags(const MessagePlan &mp, int R, bool f16 = false) {
    const int off = (int)(frags.size() / 8);
    std::vector<float> F((size_t)R * R, 0.0f);
    if (mp.rot_mode == kRotFilter) {
      const int LR = (int)mp.rot_taps.size();
      const int n = (LR - 1) / 2;
      for (int ro = 0; ro < R; ++ro)
        for (int j = 0; j < LR; ++j) F[(size_t)ro * R + (((ro - n + j) % R) + R) % R] += mp.rot_taps[j];
    } else if (mp.rot_mode == kRotCopy) {
      for (int ro = 0; ro < R; ++ro) F[(size_t)ro * R + ro] = 1.0f;
    }
    std::vector<uint16_t> lo;
    const int MT = (R + 31) / 32, KS = (R + 15) / 16;
    for (int mt = 0; mt < MT; ++mt)
      for (int ks = 0; ks < KS; ++ks)
        for (int l = 0; l < 64; ++l)
          for (int t = 0; t < 8; ++t) {
            const int ro = 32 * mt + l % 32, h = l / 32;
            const int sl = 32 * (ks / 2) + 16 * (ks % 2) + 8 * (t / 4) + 4 * h + t % 4;
            const float v = (ro < R && sl < R) ? F[(size_t)ro * R + sl] : 0.0f;
            if (f16) {  // fp16 hi + lo
              const _Float16 hi = (_Float16)v;
              frags.push_back(__builtin_bit_cast(uint16_t, hi));
              lo.push_back(__builtin_bit_cast(uint16_t, (_Float16)(v - (float)hi)));
            } else {
              const uint16_t hi = bf16_rne(v);
              frags.push_back(hi);
              lo.push_back(bf16_rne(v - bf16_float(hi)));
            }
          }
    frags.insert(frags.end(), lo.begin(), lo.end());
    return off;
  }

  int add_taps(const std::vector<float> &f, int align) {
    while (taps.size() % 16) taps.push_back(0.0f);
    int off = (int)taps.size();
    taps.insert(taps.end(), f.begin(), f.end());
    while ((taps.size() - off) % align) taps.push_back(0.0f);
    return off;
  }
  int add_message(const MessagePlan &mp, MsgJob &J) {
    const int R = (int)mp.fwd_ox.size();
    J.tpitch = t_pitch(J.W);
    J.delta = mp.delta;
    J.rot_mode = mp.rot_mode;
    J.rot_len = (int)mp.rot_taps.size();
    J.kx = mp.kx;
    J.ky = mp.ky;
    J.taps_r = add_taps(mp.rot_taps, 4);
    if (mp.od) {
      // the row pass keeps the rotation-filtered grid: kx = 0 and a unit tap
      J.od = 1;
      J.od_direct = mp.fwd_direct ? 1 : 0;
      J.th = mp.th;
      J.tw = mp.tw;
      J.kxe = mp.kx;
      J.kye = mp.ky;
      J.taps_xe = add_taps(mp.fx, 16);
      J.taps_ye = add_taps(mp.fy, 16);
      J.kx = 0;
      J.ky = 0;
      J.taps_x = add_taps(std::vector<float>{1.0f}, 16);
      J.taps_y = J.taps_x;
      for (int i = 0; i < 6; ++i) {
        J.t_fwd[i] = mp.t_fwd[i];
        J.t_back[i] = mp.t_back[i];
      }
    } else {
      J.od = 0;
      J.taps_x = add_taps(mp.fx, 16);
      J.taps_y = add_taps(mp.fy, 16);
      J.kk_x = mfma_kk(mp.kx);
      J.kk_y = mfma_kk(mp.ky);
      if (J.kk_x <= kMfmaMaxKK && J.kk_y <= kMfmaMaxKK) {
        J.frag_x = add_frags(mp.fx, J.kk_x);
        J.frag_y = add_frags(mp.fy, J.kk_y);
        J.frag_y16 = add_taps16(mp.fy, J.kk_y);
        J.frag_r = (R == 24 || R == 48) ? add_rot_frags(mp, R) : 0;
        J.frag_r16 = (R == 24 || R == 48) ? add_rot_frags(mp, R, true) : 0;
        // the row-pass output is at most (sum of x taps) (largest row sum of
        // the rotation filter) times max exp(b - M) = 1; its fp16 copy holds
        // it scaled by 2^t16_shift, that bound (with margin) near 2^14
        double sx = 0.0, sr = 0.0;
        for (float f : mp.fx) sx += std::fabs(f);
        if (mp.rot_mode == kRotFilter)
          for (float f : mp.rot_taps) sr += std::fabs(f);
        else if (mp.rot_mode == kRotCopy)
          sr = 1.0;
        const double bound = 1.01 * sx * sr;
        J.t16_shift = bound > 0.0 ? std::max(-100, std::min(100, (int)std::floor(14.0 - std::log2(bound)))) : 0;
        // the x filter's output, the rotation GEMM's fp16 operand there, is at
        // most (sum of x taps) times max exp(b - M) = 1
        J.x16_shift = sx > 0.0 ? std::max(-100, std::min(100, (int)std::floor(14.0 - std::log2(1.01 * sx)))) : 0;
      } else {
        J.kk_x = J.kk_y = 0;  // too wide for the BF16 passes: FP32 only
      }
    }
    J.shifts = (int)shifts.size();
    shifts.insert(shifts.end(), mp.fwd_ox.begin(), mp.fwd_ox.end());
    shifts.insert(shifts.end(), mp.fwd_oy.begin(), mp.fwd_oy.end());
    shifts.insert(shifts.end(), mp.inv_ox.begin(), mp.inv_ox.end());
    shifts.insert(shifts.end(), mp.inv_oy.begin(), mp.inv_oy.end());
    // the forward offsets again, indexed by output slice (source = output - delta);
    // an output slice without a source gets a row offset no grid can satisfy
    for (int ro = 0; ro < R; ++ro) {
      const int rs = ro - mp.delta;
      shifts.push_back(rs >= 0 && rs < R ? mp.fwd_ox[rs] : 0);
    }
    for (int ro = 0; ro < R; ++ro) {
      const int rs = ro - mp.delta;
      shifts.push_back(rs >= 0 && rs < R ? mp.fwd_oy[rs] : kNoSourceRow);
    }
    // the same pair packed into one word (x offset in the low, y offset in the
    // high 16 bits, clamped to +-32767, which no grid of the BF16 passes can
    // reach; no source: y offset -32768)
    for (int ro = 0; ro < R; ++ro) {
      const int rs = ro - mp.delta;
      const bool has = rs >= 0 && rs < R;
      const int ox = has ? std::max(-32767, std::min(32767, mp.fwd_ox[rs])) : 0;
      const int oy = has ? std::max(-32767, std::min(32767, mp.fwd_oy[rs])) : -32768;
      shifts.push_back((int)(((unsigned)oy << 16) | ((unsigned)ox & 0xffffu)));
    }
    // the source slice of each output slice clamped into [0, R) (an address
    // inside the grid for a slice without a source, whose loads are discarded)
    for (int ro = 0; ro < R; ++ro) shifts.push_back(std::max(0, std::min(R - 1, ro - mp.delta)));
    // a warp whose offset changes along an axis: breaks and steps, forward by
    // output slice (x break, x step, y break, y step; none for a slice
    // without a source), inverse by slice
    J.split = mp.split ? 1 : 0;
    if (mp.split) {
      const std::vector<int> *fwd[4] = {&mp.fwd_xb, &mp.fwd_xd, &mp.fwd_yb, &mp.fwd_yd};
      const int none[4] = {J.W, 0, J.H, 0};
      for (int k = 0; k < 4; ++k)
        for (int ro = 0; ro < R; ++ro) {
          const int rs = ro - mp.delta;
          shifts.push_back(rs >= 0 && rs < R ? (*fwd[k])[rs] : none[k]);
        }
      for (const std::vector<int> *v : {&mp.inv_xb, &mp.inv_xd, &mp.inv_yb, &mp.inv_yd})
        shifts.insert(shifts.end(), v->begin(), v->end());
      J.kk_x = J.kk_y = 0;  // the BF16 passes take uniform warps only: FP32
    }
    return PS_OK;
  }
};

UnaryMask part_mask(const Model &m, int p, int W) {
  UnaryMask k;
  k.rot_bits = m.is_upright[p] ? upright_masked_bits(m.ep) : 0ull;
  k.strip = 0;
  k.pad_ = 0;
  if (p == m.tree.root && m.ep.strip_border_detections > 0) k.strip = (int)(m.ep.strip_border_detections * W);
  return k;
}

}  // namespace

extern "C" int ps_plan_create(const ps_model *model, const ps_instance *instances, int n_instances, ps_plan **out) {
  if (!out) return set_error(PS_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!model || !instances || n_instances < 1) return set_error(PS_ERR_INVALID_ARG, "model, instances and n_instances >= 1 required");
  const Model &m = model->m;
  const Tree &t = m.tree;
  const int P = m.P, S = (int)m.ep.num_scale_steps, R = (int)m.ep.num_rotation_steps;
  const int K = (int)t.root_joints.size();
  for (int i = 0; i < n_instances; ++i) {
    const ps_instance &in = instances[i];
    if (in.height < 1 || in.width < 1) return set_error(PS_ERR_INVALID_ARG, "instance sizes must be >= 1");
    if ((long long)R * in.height * t_pitch(in.width) >= (1ll << 31))
      return set_error(PS_ERR_UNSUPPORTED, "R*H*W (rows padded to 32) must be below 2^31");
    if (in.unary_offset < 0 || in.marginal_offset < 0 || in.root_offset < 0)
      return set_error(PS_ERR_INVALID_ARG, "negative offset");
    if (in.joints) {
      int st = check_joints(m, in.joints);
      if (st != PS_OK) return st;
    }
  }
  int device = 0;
  int st = check_device(device);
  if (st != PS_OK) return st;

  // chains: nodes from the root child down to the leaf
  std::vector<std::vector<int>> chains(K);
  for (int k = 0; k < K; ++k) {
    int p = m.joints[t.root_joints[k]].child_idx;
    while (p >= 0) {
      chains[k].push_back(p);
      int cj = t.child_joint[p];
      p = cj >= 0 ? m.joints[cj].child_idx : -1;
    }
  }

  // inner parts (a parent and a child): each keeps its upward belief in a
  // grid of its own until its downward message
  std::vector<int> inner_idx(P, -1);
  int n_inner = 0;
  for (int p = 0; p < P; ++p)
    if (p != t.root && t.child_joint[p] >= 0) inner_idx[p] = n_inner++;

  Builder B(m);
  const int items = n_instances * S;
  std::vector<std::vector<MsgJob>> up_jobs(t.max_height), dn_jobs(t.max_depth);
  size_t scratch_off = 0;
  std::vector<long long> keep_offsets(n_instances);
  long long keep_total = 0;
  int max_rows = 0, max_hw = 0;
  for (int i = 0; i < n_instances; ++i) {
    const ps_instance &in = instances[i];
    const int H = in.height, W = in.width;
    const long long n = (long long)R * H * W;
    const ps_joint *J = in.joints ? in.joints : m.joints.data();
    max_rows = std::max(max_rows, R * H);
    max_hw = std::max(max_hw, H * W);
    const int kty = (H + 63) / 64, ktx = (W + 63) / 64;
    keep_offsets[i] = keep_total;
    keep_total += (long long)P * S * R * kty * ktx;
    for (int s = 0; s < S; ++s) {
      const int it = i * S + s;
      const double scale = scale_from_index(m.ep, s);
      // the caller's grids are dense; the plan's own start every row on a
      // 128-byte boundary (and every grid on a 256-byte one)
      const int gp = grid_pitch(W);
      auto unary = [&](int p) { return gref(kBufUnary, in.unary_offset + ((long long)p * S + s) * n, W); };
      auto marg = [&](int p) { return gref(kBufMarginal, in.marginal_offset + ((long long)p * S + s) * n, W); };
      const long long np64 = ((long long)R * H * gp + 63) & ~63ll;
      const long long nt64 = ((long long)R * H * t_pitch(W) + 63) & ~63ll;  // row-pass outputs
      auto tbuf = [&](int k) { return gref(kBufScratch, (long long)scratch_off + (long long)k * nt64, t_pitch(W)); };
      // message slot of chain k: the upward message into the root, then the
      // downward sources of the chain past its first message
      auto mbuf = [&](int k) { return gref(kBufScratch, (long long)scratch_off + K * nt64 + (long long)k * np64, gp); };
      // upward belief of an inner part (a part with a parent and a child),
      // kept from its upward message to its downward one
      auto bel = [&](int p) {
        return gref(kBufScratch, (long long)scratch_off + K * nt64 + (K + inner_idx[p]) * np64, gp);
      };
      // log_from_root of chain k: written by the root combine into the
      // marginal grid of the chain's first part (dense, so the root writes all
      // its grids along its pixels), read by the row pass of the chain's first
      // downward message, whose column pass then writes that part's marginal
      // over it
      auto lfr = [&](int k) { return marg(chains[k][0]); };
      auto up_key = [&](int p) { return it * 2 * P + p; };
      auto dn_key = [&](int p) { return it * 2 * P + P + p; };
      auto slot = [&](int p) { return it * P + p; };
      auto keep_plane = [&](MsgJob &mj, int p) {
        mj.keep_off = keep_offsets[i] + ((long long)p * S + s) * R * kty * ktx;
        mj.keep_ty = kty;
        mj.keep_tx = ktx;
      };

      // leaf maxima
      for (int k = 0; k < K; ++k) {
        int leaf = chains[k].back();
        MaxJob mj;
        std::memset(&mj, 0, sizeof(mj));
        mj.H = H;
        mj.W = W;
        mj.inst = i;
        mj.src_kind = m.is_detect[leaf] ? kSrcUnary : kSrcZero;
        mj.key = up_key(leaf);
        mj.src = unary(leaf);
        mj.mask = part_mask(m, leaf, W);
        B.maxes.push_back(mj);
      }
      // message plans of every edge: upward with the model's sparse flag
      // (objectdetect_findrot.cpp:630-635), downward with offsets swapped,
      // -rot_mean and bilinear rotations (:174-179, :224-229)
      std::vector<std::vector<MessagePlan>> upp(K), dnp(K);
      std::vector<long long> od_floats(K, 0);
      for (int k = 0; k < K; ++k) {
        const int len = (int)chains[k].size();
        upp[k].resize(len);
        dnp[k].resize(len);
        for (int d = 1; d <= len; ++d) {
          const int jidx = t.parent_joint[chains[k][d - 1]];
          const ps_joint &jt = J[jidx];
          const std::string where = "instance " + std::to_string(i) + ", joint " + std::to_string(jidx);
          st = build_message_plan(m.ep, H, W, jt.offset_c, jt.offset_p, jt.C, jt.rot_mean, jt.rot_sigma, scale,
                                  m.is_sparse != 0, upp[k][d - 1]);
          if (st != PS_OK) return set_error(st, where + " (up): " + g_last_error);
          st = build_message_plan(m.ep, H, W, jt.offset_p, jt.offset_c, jt.C, -jt.rot_mean, jt.rot_sigma, scale,
                                  false, dnp[k][d - 1]);
          if (st != PS_OK) return set_error(st, where + " (down): " + g_last_error);
          for (const MessagePlan *mp : {&upp[k][d - 1], &dnp[k][d - 1]})
            if (mp->od) od_floats[k] = std::max(od_floats[k], (od_scratch_floats(R, mp->th, mp->tw) + 63) & ~63ll);
        }
      }
      // eigenbasis scratch of chain k, after the T grids and message slots
      std::vector<long long> od_off(K, 0);
      long long od_total = 0;
      for (int k = 0; k < K; ++k) {
        od_off[k] = (long long)scratch_off + K * nt64 + (K + n_inner) * np64 + od_total;
        od_total += od_floats[k];
      }
      auto odbuf = [&](int k) { return gref(kBufScratch, od_off[k], 0); };

      // upward levels, each message as late as its parent allows: the edge
      // into a node at depth d runs at level max_height - d, so every message
      // into the root (a one-part chain's too) runs in the last level and a
      // level's messages share one destination kind in the person tree
      for (int k = 0; k < K; ++k) {
        const int len = (int)chains[k].size();
        for (int d = len; d >= 1; --d) {
          const int c = chains[k][d - 1];
          const int jidx = t.parent_joint[c];
          const int p = m.joints[jidx].parent_idx;
          const MessagePlan &mp = upp[k][d - 1];
          MsgJob mj;
          std::memset(&mj, 0, sizeof(mj));
          mj.H = H;
          mj.W = W;
          mj.inst = i;
          B.add_message(mp, mj);
          mj.od_buf = mp.od ? odbuf(k) : kNone;
          if (t.height[c] == 1) {
            mj.src_kind = m.is_detect[c] ? kSrcUnary : kSrcZero;
            mj.src = unary(c);
            mj.src_mask = part_mask(m, c, W);
          } else {
            mj.src_kind = kSrcGrid;
            mj.src = bel(c);
          }
          mj.src_key = up_key(c);
          mj.tmp = tbuf(k);
          mj.dst2 = kNone;
          mj.dst_bel = kNone;
          mj.dst_unary = kNone;
          keep_plane(mj, p);
          if (p == t.root) {
            mj.dst_kind = kDstUpRoot;
            mj.dst = mbuf(k);
          } else {
            mj.dst_kind = kDstUpNode;
            mj.dst = bel(p);
            mj.dst_unary = unary(p);
            mj.dst_mask = part_mask(m, p, W);
            mj.dst_add_unary = m.is_detect[p];
            mj.dst_key = up_key(p);
          }
          up_jobs[t.max_height - t.depth[c]].push_back(mj);
        }
      }
      // root combine
      {
        RootJob rj;
        std::memset(&rj, 0, sizeof(rj));
        rj.H = H;
        rj.W = W;
        rj.inst = i;
        rj.K = K;
        rj.add_unary = m.is_detect[t.root];
        for (int k = 0; k < K; ++k) {
          rj.msg[k] = mbuf(k);
          rj.fr[k] = lfr(k);
          rj.fr_key[k] = dn_key(chains[k][0]);
        }
        rj.unary = unary(t.root);
        rj.mask = part_mask(m, t.root, W);
        rj.marginal = marg(t.root);
        rj.map_slot = slot(t.root);
        rj.root_post = gref(kBufRoot, in.root_offset + (long long)s * H * W, W);
        if (m.is_upright[t.root]) {
          int k1 = index_from_rot(m.ep, -1e-6), k2 = index_from_rot(m.ep, 1e-6);
          if (k1 < 0 || k2 < 0 || k1 >= R || k2 >= R)
            return set_error(PS_ERR_INVALID_ARG, "upright root needs rotation 0 inside the rotation range");
          rj.valid_rot = (1ull << k1) | (1ull << k2);
        } else {
          rj.valid_rot = R == 64 ? ~0ull : ((1ull << R) - 1);
        }
        B.roots.push_back(rj);
      }
      // downward levels, each message as late as its subtree allows: the edge
      // into a node of height h runs at level max_depth - h (every leaf's in
      // the last level); its parent's message ran in an earlier one
      for (int k = 0; k < K; ++k) {
        const int len = (int)chains[k].size();
        for (int d = 1; d <= len; ++d) {
          const int c = chains[k][d - 1];
          const MessagePlan &mp = dnp[k][d - 1];
          MsgJob mj;
          std::memset(&mj, 0, sizeof(mj));
          mj.H = H;
          mj.W = W;
          mj.inst = i;
          B.add_message(mp, mj);
          mj.od_buf = mp.od ? odbuf(k) : kNone;
          mj.src_kind = kSrcGrid;
          mj.src = d == 1 ? lfr(k) : mbuf(k);
          mj.src_key = dn_key(c);
          mj.tmp = tbuf(k);
          mj.dst_kind = kDstDown;
          mj.dst = marg(c);
          keep_plane(mj, c);
          mj.dst_unary = unary(c);
          mj.dst_mask = part_mask(m, c, W);
          const bool has_child = t.child_joint[c] >= 0;
          mj.dst_base = has_child ? 0 : (m.is_detect[c] ? 1 : 2);
          mj.dst_bel = kNone;
          if (has_child) {
            mj.dst_bel = bel(c);
            mj.dst2 = mbuf(k);
            mj.dst2_key = dn_key(m.joints[t.child_joint[c]].child_idx);
          } else {
            mj.dst2 = kNone;
          }
          mj.map_slot = slot(c);
          dn_jobs[t.max_depth - t.height[c]].push_back(mj);
        }
      }
      for (int p = 0; p < P; ++p) {
        MapJob mj;
        std::memset(&mj, 0, sizeof(mj));
        mj.H = H;
        mj.W = W;
        mj.inst = i;
        mj.slot = slot(p);
        mj.out = it * P + p;
        B.maps.push_back(mj);
      }
      scratch_off += (size_t)(K * nt64 + (K + n_inner) * np64 + od_total);
    }
  }

  ps_plan *pl = new ps_plan();
  pl->model = m;
  pl->device = device;
  pl->n_inst = n_instances;
  pl->S = S;
  pl->P = P;
  pl->R = R;
  pl->items = items;
  pl->max_rows = max_rows;
  pl->max_hw = max_hw;
  pl->keep_offsets = keep_offsets;
  pl->keep_bytes = keep_total;
  auto add_levels = [&](std::vector<std::vector<MsgJob>> &lv, std::vector<Level> &dst) -> int {
    for (auto &v : lv) {
      std::stable_partition(v.begin(), v.end(), [](const MsgJob &j) { return j.od == 0; });
      // diagonal jobs by width class (the order of a level's jobs is free)
      auto mid = std::find_if(v.begin(), v.end(), [](const MsgJob &j) { return j.od != 0; });
      // and, within a width class, by epilogue kind: the column pass launches
      // each kind on its own, so a kind that needs fewer registers runs at
      // its own occupancy and every block of a launch does the same work
      std::stable_sort(v.begin(), mid, [](const MsgJob &a, const MsgJob &b2) {
        const int ca = (a.W + 63) / 64, cb = (b2.W + 63) / 64;
        return ca != cb ? ca < cb : epi_kind(a) < epi_kind(b2);
      });
      Level L;
      auto add_to = [](std::vector<WidthClass> &cl, int i, const MsgJob &j, bool split_kind) {
        if (cl.empty() || (cl.back().max_w + 63) / 64 != (j.W + 63) / 64 ||
            (split_kind && cl.back().epi_hi != epi_kind(j))) {
          WidthClass c;
          c.start = i;
          cl.push_back(c);
        }
        WidthClass &c = cl.back();
        ++c.count;
        c.max_h = std::max(c.max_h, j.H);
        c.max_w = std::max(c.max_w, j.W);
        c.epi_hi = std::max(c.epi_hi, epi_kind(j));
      };
      for (auto it = v.begin(); it != mid; ++it) {
        const int i = (int)(it - v.begin());
        add_to(L.classes, i, *it, false);
        add_to(L.col_classes, i, *it, true);
      }
      for (auto &c : L.classes) class_runs(c, v);
      for (auto &c : L.col_classes) class_runs(c, v);
      L.off = (int)B.jobs.size();
      L.count = (int)v.size();
      L.od.start = (int)(mid - v.begin());
      L.od.count = (int)(v.end() - mid);
      class_runs(L.od, v);
      L.mfma_ok = R == 24 || R == 48;
      for (MsgJob &j : v) {
        if (j.split) L.split = true;
        L.max_h = std::max(L.max_h, j.H);
        L.max_w = std::max(L.max_w, j.W);
        L.lds = std::max(L.lds, rowpass_lds_bytes(R, j.W, j.kx));
        if (j.od) {
          L.od_cells = std::max(L.od_cells, (long long)R * j.th * j.tw);
          L.od_hw = std::max(L.od_hw, (long long)j.H * j.W);
        } else {
          ++L.n_diag;
          L.col_lds = std::max(L.col_lds, colpass_lds_bytes(j.ky, 16));
          // (the BF16 passes address a row-pass output grid with 32-bit byte offsets)
          if (j.kk_x < 1 || j.kk_y < 1 || j.W > 512 || j.H > 32767 || 4ll * R * j.H * j.tpitch >= (1ll << 32))
            L.mfma_ok = false;
          if (L.mfma_ok) {
            for (int t = 0; t < 3; ++t) {
              L.mfma_row_lds[t] = std::max(L.mfma_row_lds[t], rowpass_mfma_lds_bytes(R, j.W, j.kx, j.kk_x, t + 1));
              L.mfma_col_lds[t] = std::max(L.mfma_col_lds[t], colpass_mfma_lds_bytes(j.kk_y, t + 1));
            }
            L.mfma_col16_lds = std::max(L.mfma_col16_lds, colpass_f16_lds_bytes(j.kk_y, j.H));
          }
        }
      }
      if (L.mfma_row_lds[2] > kLdsLimit || L.mfma_col_lds[2] > kLdsLimit || L.mfma_col16_lds > kLdsLimit)
        L.mfma_ok = false;
      if (L.lds > kLdsLimit || L.col_lds > kLdsLimit)
        return set_error(PS_ERR_UNSUPPORTED, "row pass needs " + std::to_string(L.lds) +
                                                 " bytes of LDS (R * width too large for one workgroup)");
      B.jobs.insert(B.jobs.end(), v.begin(), v.end());
      dst.push_back(L);
    }
    return PS_OK;
  };
  st = add_levels(up_jobs, pl->up);
  if (st == PS_OK) st = add_levels(dn_jobs, pl->down);
  // the kernels share one byte offset between grids of equal row pitch: the
  // epilogues between dst and the unary a marginal adds, and between the next
  // source and the stored belief; the root between its message slots, and
  // between the unary and the marginal (dense, like the root posterior)
  if (st == PS_OK) {
    for (const MsgJob &j : B.jobs) {
      const int kind = epi_kind(j);
      bool ok = j.src.pitch >= j.W && j.dst.pitch >= j.W;
      if (kind >= kEpiDownLeaf) ok = ok && j.dst_unary.pitch == j.dst.pitch;
      if (kind == kEpiDownMid) ok = ok && j.dst_bel.pitch == j.dst2.pitch;
      if (!ok) st = set_error(PS_ERR_INVALID_ARG, "internal: grid pitches of a message do not pair up");
    }
    for (const RootJob &j : B.roots) {
      bool ok = j.unary.pitch == j.W && j.marginal.pitch == j.W;
      for (int k = 0; k < j.K; ++k)
        ok = ok && j.msg[k].pitch == j.msg[0].pitch && j.msg[k].pitch >= j.W && j.fr[k].pitch == j.W;
      if (!ok) st = set_error(PS_ERR_INVALID_ARG, "internal: grid pitches of a root do not pair up");
    }
  }
  if (st != PS_OK) {
    delete pl;
    return st;
  }
  pl->n_roots = (int)B.roots.size();
  pl->root_children = B.roots.empty() ? 0 : B.roots[0].K;
  pl->n_maxes = (int)B.maxes.size();
  pl->all_mfma = true;
  for (const auto *lv : {&pl->up, &pl->down})
    for (const Level &L : *lv)
      if (!L.mfma_ok || L.n_diag != L.count) pl->all_mfma = false;
  pl->max_children = K;
  pl->n_keys = items * 2 * P;
  pl->n_map = items * P;
  pl->scratch_elems = scratch_off;

  // one allocation for the tables, one for the scratch grids
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t o_jobs = 0;
  size_t o_roots = align(o_jobs + B.jobs.size() * sizeof(MsgJob));
  size_t o_maxes = align(o_roots + B.roots.size() * sizeof(RootJob));
  size_t o_maps = align(o_maxes + B.maxes.size() * sizeof(MaxJob));
  size_t o_taps = align(o_maps + B.maps.size() * sizeof(MapJob));
  size_t o_shifts = align(o_taps + B.taps.size() * sizeof(float));
  size_t o_frags = align(o_shifts + B.shifts.size() * sizeof(int));
  size_t o_keys = align(o_frags + B.frags.size() * sizeof(uint16_t));
  size_t o_mkeys = align(o_keys + (size_t)pl->n_keys * sizeof(unsigned));
  size_t o_top2 = align(o_mkeys + (size_t)pl->n_map * sizeof(unsigned long long));
  size_t o_flags = align(o_top2 + (size_t)pl->n_map * sizeof(unsigned long long));
  size_t o_nfb = align(o_flags + (size_t)n_instances * sizeof(unsigned));
  size_t total = align(o_nfb + (size_t)n_instances * sizeof(unsigned));
  hipError_t e = hipMalloc(&pl->tables, total);
  if (e != hipSuccess) {
    delete pl;
    return set_error(PS_ERR_OUT_OF_MEMORY, "tables: " + std::string(hipGetErrorString(e)));
  }
  char *base = static_cast<char *>(pl->tables);
  pl->jobs = reinterpret_cast<MsgJob *>(base + o_jobs);
  pl->roots = reinterpret_cast<RootJob *>(base + o_roots);
  pl->maxes = reinterpret_cast<MaxJob *>(base + o_maxes);
  pl->maps = reinterpret_cast<MapJob *>(base + o_maps);
  pl->taps = reinterpret_cast<float *>(base + o_taps);
  pl->shifts = reinterpret_cast<int *>(base + o_shifts);
  pl->frags = reinterpret_cast<uint4 *>(base + o_frags);
  pl->keys = reinterpret_cast<unsigned *>(base + o_keys);
  pl->map_keys = reinterpret_cast<unsigned long long *>(base + o_mkeys);
  pl->top2 = reinterpret_cast<unsigned long long *>(base + o_top2);
  pl->flags = reinterpret_cast<unsigned *>(base + o_flags);
  pl->not_fallback = reinterpret_cast<unsigned *>(base + o_nfb);
  std::vector<char> host(total, 0);
  std::memcpy(host.data() + o_jobs, B.jobs.data(), B.jobs.size() * sizeof(MsgJob));
  std::memcpy(host.data() + o_roots, B.roots.data(), B.roots.size() * sizeof(RootJob));
  std::memcpy(host.data() + o_maxes, B.maxes.data(), B.maxes.size() * sizeof(MaxJob));
  std::memcpy(host.data() + o_maps, B.maps.data(), B.maps.size() * sizeof(MapJob));
  std::memcpy(host.data() + o_taps, B.taps.data(), B.taps.size() * sizeof(float));
  std::memcpy(host.data() + o_shifts, B.shifts.data(), B.shifts.size() * sizeof(int));
  std::memcpy(host.data() + o_frags, B.frags.data(), B.frags.size() * sizeof(uint16_t));
  e = hipMemcpy(pl->tables, host.data(), total, hipMemcpyHostToDevice);
  if (e == hipSuccess && pl->scratch_elems) {
    e = hipMalloc(&pl->scratch, pl->scratch_elems * sizeof(float));
    if (e != hipSuccess) {
      (void)hipFree(pl->tables);
      delete pl;
      return set_error(PS_ERR_OUT_OF_MEMORY, "scratch (" + std::to_string(pl->scratch_elems * 4) + " bytes): " + hipGetErrorString(e));
    }
  }
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&pl->flags_host), sizeof(unsigned) * n_instances);
  // room for every job twice (row-pass and column-pass classes list the same
  // jobs) and every root, max and map job, each part 256-byte aligned
  pl->cbytes = align(2 * B.jobs.size() * sizeof(MsgJob)) + align(B.roots.size() * sizeof(RootJob)) +
               align(B.maxes.size() * sizeof(MaxJob)) + align(B.maps.size() * sizeof(MapJob));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&pl->cstage), pl->cbytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&pl->ctables), pl->cbytes);
  if (e != hipSuccess) {
    if (pl->ctables) (void)hipFree(pl->ctables);
    if (pl->cstage) (void)hipHostFree(pl->cstage);
    if (pl->flags_host) (void)hipHostFree(pl->flags_host);
    if (pl->scratch) (void)hipFree(pl->scratch);
    (void)hipFree(pl->tables);
    delete pl;
    return hip_error(e, "plan upload");
  }
  pl->h_jobs = B.jobs;
  pl->h_roots = B.roots;
  pl->h_maxes = B.maxes;
  pl->h_maps = B.maps;
  size_t max_lds = 0, max_col_lds = 0;
  for (auto *lv : {&pl->up, &pl->down})
    for (auto &L : *lv) {
      max_lds = std::max(max_lds, L.lds);
      max_col_lds = std::max(max_col_lds, L.col_lds);
    }
  e = rowpass_set_lds_limit(max_lds);
  if (e == hipSuccess) e = colpass_set_lds_limit(max_col_lds);
  size_t mrow = 0, mcol = 0;
  for (auto *lv : {&pl->up, &pl->down})
    for (auto &L : *lv)
      if (L.mfma_ok) {
        mrow = std::max(mrow, L.mfma_row_lds[2]);
        mcol = std::max(mcol, std::max(L.mfma_col_lds[2], L.mfma_col16_lds));
      }
  if (e == hipSuccess) e = mfma_set_lds_limit(mrow, mcol);
  if (e != hipSuccess) {
    ps_plan_destroy(pl);
    return hip_error(e, "hipFuncSetAttribute");
  }
  *out = pl;
  return PS_OK;
}

extern "C" int ps_plan_set_method(ps_plan *plan, int method) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  int terms, col_terms;
  switch (method) {
    case PS_METHOD_FP32: terms = col_terms = 0; break;
    case PS_METHOD_BF16: terms = col_terms = 1; break;
    case PS_METHOD_BF16X2: terms = col_terms = 2; break;
    case PS_METHOD_BF16X3: terms = col_terms = 3; break;
    case PS_METHOD_BF16_MIXED: terms = 1; col_terms = 3; break;
    default: return set_error(PS_ERR_INVALID_ARG, "unknown method " + std::to_string(method));
  }
  if ((terms != plan->terms || col_terms != plan->col_terms) && plan->gexec) {
    (void)hipGraphExecDestroy(plan->gexec);
    plan->gexec = nullptr;
  }
  plan->terms = terms;
  plan->col_terms = col_terms;
  return PS_OK;
}

extern "C" int ps_plan_set_map_guard(ps_plan *plan, float tau) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  if (!(tau >= 0.0f) || !std::isfinite(tau)) return set_error(PS_ERR_INVALID_ARG, "tau must be finite and >= 0");
  if (tau != plan->guard_tau && plan->gexec) {
    (void)hipGraphExecDestroy(plan->gexec);
    plan->gexec = nullptr;
  }
  plan->guard_tau = tau;
  return PS_OK;
}

extern "C" int ps_plan_read_map_guard(ps_plan *plan, int *n_reevaluated, unsigned *flags_host, void *stream) {
  if (!plan || !n_reevaluated) return set_error(PS_ERR_INVALID_ARG, "null argument");
  std::vector<unsigned> f(plan->n_inst);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(f.data(), plan->flags, sizeof(unsigned) * f.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_error(e, "reading the MAP guard flags");
  int n = 0;
  for (int i = 0; i < plan->n_inst; ++i) {
    n += f[i] ? 1 : 0;
    if (flags_host) flags_host[i] = f[i];
  }
  *n_reevaluated = n;
  return PS_OK;
}

namespace {
// the row-pass launch classes of a plan in execution order
std::vector<WidthClass *> tune_classes(ps_plan *plan) {
  std::vector<WidthClass *> v;
  for (auto *lv : {&plan->up, &plan->down})
    for (auto &L : *lv)
      for (auto &c : L.classes) v.push_back(&c);
  return v;
}
int check_knob(int knob, int value) {
  switch (knob) {
    case PS_TUNE_MFMA_ROWS:
      if (value < 1 || value > 64) return set_error(PS_ERR_INVALID_ARG, "rows per block must be in [1, 64]");
      return PS_OK;
    case PS_TUNE_FP32_ROW_THREADS:
      if (value < 0 || value > 512 || value % 64)
        return set_error(PS_ERR_INVALID_ARG, "threads per block must be 0 or a multiple of 64 up to 512");
      return PS_OK;
    case PS_TUNE_FP32_COL_ROWS:
      if (value != 0 && value != 8 && value != 16)
        return set_error(PS_ERR_INVALID_ARG, "column-pass rows per lane must be 0, 8 or 16");
      return PS_OK;
    default:
      return set_error(PS_ERR_INVALID_ARG, "unknown tuning knob " + std::to_string(knob));
  }
}
// the column-pass classes a row-pass class's knob reaches: those of its
// level with its width class (64-column strips)
std::vector<WidthClass *> col_classes_of(ps_plan *plan, int cls) {
  std::vector<WidthClass *> out;
  int k = 0;
  for (auto *lv : {&plan->up, &plan->down})
    for (auto &L : *lv)
      for (auto &c : L.classes) {
        if (k++ != cls) continue;
        const int xt = (c.max_w + 63) / 64;
        for (auto &cc : L.col_classes)
          if ((cc.max_w + 63) / 64 == xt) out.push_back(&cc);
      }
  return out;
}
void set_knob(ps_plan *plan, int cls, WidthClass &c, int knob, int value) {
  if (knob == PS_TUNE_MFMA_ROWS) c.mfma_rows = value;
  else if (knob == PS_TUNE_FP32_ROW_THREADS) c.row_threads = value;
  else
    for (WidthClass *cc : col_classes_of(plan, cls)) cc->col_rows = value;
}
void drop_graph(ps_plan *plan) {
  if (plan->gexec) {
    (void)hipGraphExecDestroy(plan->gexec);
    plan->gexec = nullptr;
  }
}
}  // namespace

extern "C" int ps_plan_tune(ps_plan *plan, int knob, int value) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  int st = check_knob(knob, value);
  if (st != PS_OK) return st;
  drop_graph(plan);
  if (knob == PS_TUNE_MFMA_ROWS) plan->mfma_rows = value;
  std::vector<WidthClass *> v = tune_classes(plan);
  for (size_t k = 0; k < v.size(); ++k) set_knob(plan, (int)k, *v[k], knob, knob == PS_TUNE_MFMA_ROWS ? 0 : value);
  return PS_OK;
}

extern "C" int ps_plan_tune_class(ps_plan *plan, int knob, int cls, int value) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  int st = check_knob(knob, value);
  if (st != PS_OK) return st;
  std::vector<WidthClass *> v = tune_classes(plan);
  if (cls < 0 || cls >= (int)v.size()) return set_error(PS_ERR_INVALID_ARG, "no launch class " + std::to_string(cls));
  drop_graph(plan);
  set_knob(plan, cls, *v[cls], knob, value);
  return PS_OK;
}

extern "C" int ps_plan_classes(const ps_plan *plan, int *n_classes, int *info, int max_classes) {
  if (!plan || !n_classes) return set_error(PS_ERR_INVALID_ARG, "null argument");
  int n = 0, level = 0;
  for (auto *lv : {&plan->up, &plan->down})
    for (auto &L : *lv) {
      for (auto &c : L.classes) {
        if (info && n < max_classes) {
          info[4 * n] = level;
          info[4 * n + 1] = c.count;
          info[4 * n + 2] = c.max_h;
          info[4 * n + 3] = c.max_w;
        }
        ++n;
      }
      ++level;
    }
  *n_classes = n;
  return PS_OK;
}

extern "C" int ps_plan_method_levels(const ps_plan *plan, int *levels, int *mfma_levels) {
  if (!plan || !levels || !mfma_levels) return set_error(PS_ERR_INVALID_ARG, "null argument");
  *levels = (int)(plan->up.size() + plan->down.size());
  int n = 0;
  for (auto *lv : {&plan->up, &plan->down})
    for (auto &L : *lv) n += L.mfma_ok ? 1 : 0;
  *mfma_levels = n;
  return PS_OK;
}

extern "C" int ps_plan_set_profiling(ps_plan *plan, int enable) {
  return ps_plan_set_profiling_classes(plan, enable ? (1u << kProfileClasses) - 1u : 0u);
}

extern "C" int ps_plan_set_profiling_classes(ps_plan *plan, unsigned mask) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  if (mask >> kProfileClasses) return set_error(PS_ERR_INVALID_ARG, "profile mask has bits above the kernel classes");
  plan->profile_mask = mask;
  return PS_OK;
}

extern "C" int ps_plan_read_profile(ps_plan *plan, double ms[6], int launches[6]) {
  if (!plan || !ms || !launches) return set_error(PS_ERR_INVALID_ARG, "null argument");
  for (int c = 0; c < kProfileClasses; ++c) {
    ms[c] = 0.0;
    launches[c] = 0;
  }
  for (size_t i = 0; i < plan->n_recorded; ++i) {
    hipError_t e = hipEventSynchronize(plan->events[2 * i + 1]);
    if (e != hipSuccess) return hip_error(e, "hipEventSynchronize");
    float t = 0.0f;
    e = hipEventElapsedTime(&t, plan->events[2 * i], plan->events[2 * i + 1]);
    if (e != hipSuccess) return hip_error(e, "hipEventElapsedTime");
    ms[plan->event_class[i]] += t;
    launches[plan->event_class[i]] += 1;
  }
  plan->n_recorded = 0;
  return PS_OK;
}

extern "C" void ps_plan_destroy(ps_plan *plan) {
  if (!plan) return;
  for (hipEvent_t e : plan->events) (void)hipEventDestroy(e);
  if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
  if (plan->scratch) (void)hipFree(plan->scratch);
  if (plan->tables) (void)hipFree(plan->tables);
  if (plan->flags_host) (void)hipHostFree(plan->flags_host);
  if (plan->cstage) (void)hipHostFree(plan->cstage);
  if (plan->ctables) (void)hipFree(plan->ctables);
  delete plan;
}

extern "C" size_t ps_plan_scratch_bytes(const ps_plan *plan) {
  return plan ? plan->scratch_elems * sizeof(float) : 0;
}

// One sweep of the plan with the spatial passes of `terms` / `col_terms`
// (0: FP32) over the instances b.active names (all when null).
static hipError_t sweep(ps_plan *pl, const Bases &b, int terms, int col_terms, ps_map *map_dev, hipStream_t s,
                        ps_plan *pp, unsigned *clear_flags) {
  long ev = profile_begin(pp, 0, s);
  hipError_t e = launch_reset(b, pl->n_keys, pl->n_map, pl->n_inst, clear_flags, pl->n_inst, s);
  profile_end(pp, ev, s);
  if (e == hipSuccess && !((terms == 1 || terms == 2) && pl->all_mfma)) {
    ev = profile_begin(pp, 1, s);
    e = launch_max(b, pl->maxes, pl->n_maxes, pl->R, pl->max_hw, s);
    profile_end(pp, ev, s);
  }
  auto level = [&](const Level &L) {
    const bool mf = terms != 0 && L.mfma_ok;
    const int ti = terms > 0 ? terms - 1 : 0;
    const int tc = col_terms > 0 ? col_terms - 1 : 0;
    const MsgJob *jobs = pl->jobs + L.off;
    const int n_od = L.count - L.n_diag;
    // every kernel launch (the eigenbasis stages of a level as one) has its
    // own pair of profiling events
    for (const WidthClass &c : L.classes) {
      if (e != hipSuccess) break;
      const long ev2 = profile_begin(pp, 2, s);
      if (mf)
        e = launch_rowpass_mfma(b, jobs + c.start, c.count, pl->taps, pl->shifts, pl->frags, pl->R, terms,
                                c.mfma_rows > 0 ? c.mfma_rows : pl->mfma_rows, c.max_h, c.max_w,
                                L.mfma_row_lds[ti], c.leaf, s);
      else
        e = launch_rowpass(b, jobs + c.start, c.count, pl->taps, pl->shifts, pl->R, c.max_h, c.max_w, L.lds, s,
                           c.row_threads, c.split);
      profile_end(pp, ev2, s);
    }
    if (e == hipSuccess && n_od > 0) {  // off-diagonal messages: the FP32 row pass, then the eigenbasis stages
      const long ev2 = profile_begin(pp, 2, s);
      e = launch_rowpass(b, jobs + L.n_diag, n_od, pl->taps, pl->shifts, pl->R, L.max_h, L.max_w, L.lds, s, 0,
                         L.split);
      profile_end(pp, ev2, s);
    }
    if (e != hipSuccess) return;
    // the fp16 and FP32 column passes have one instantiation per highest
    // epilogue kind: they launch per (width, kind) class; the bf16 three-term
    // column pass per width
    const bool per_kind = !mf || terms != 3;
    for (const WidthClass &c : per_kind ? L.col_classes : L.classes) {
      if (e != hipSuccess) break;
      const long ev2 = profile_begin(pp, 3, s);
      if (mf)
        e = launch_colpass_mfma(b, jobs + c.start, c.count, pl->frags, pl->shifts, pl->R, col_terms, terms != 3,
                                c.epi_hi, c.max_w, terms != 3 ? L.mfma_col16_lds : L.mfma_col_lds[tc], c.leaf, s);
      else
        e = launch_colpass(b, jobs + c.start, c.count, pl->taps, pl->shifts, pl->R, c.epi_hi, c.max_w, L.col_lds, s,
                           c.col_rows, c.split);
      profile_end(pp, ev2, s);
    }
    // the column pass class also times the eigenbasis stages
    if (e == hipSuccess && n_od > 0) {
      const long ev2 = profile_begin(pp, 3, s);
      e = launch_offdiag(b, jobs + L.n_diag, n_od, pl->taps, pl->shifts, pl->R, L.od_cells, L.od_hw, s);
      profile_end(pp, ev2, s);
    }
  };
  for (const Level &L : pl->up) {
    if (e != hipSuccess) break;
    level(L);
  }
  if (e == hipSuccess) {
    ev = profile_begin(pp, 4, s);
    // the BF16 methods accumulate the root's log-sum-exp in FP64 (shifted),
    // the FP32 sweep in FP32 as the reference does
    e = launch_root(b, pl->roots, pl->n_roots, pl->root_children, pl->R, pl->max_hw, terms != 0, s);
    profile_end(pp, ev, s);
  }
  for (const Level &L : pl->down) {
    if (e != hipSuccess) break;
    level(L);
  }
  if (e == hipSuccess && map_dev) {
    ev = profile_begin(pp, 5, s);
    e = launch_map_decode(b, pl->maps, pl->n_map, map_dev, s);
    profile_end(pp, ev, s);
  }
  return e;
}

// The FP32 sweep of the instances whose flag is set in pl->flags_host (and,
// on the device, in b.active): their jobs are gathered into compact tables
// (one segment per launch class, uploaded from pinned memory in stream
// order), so every launch covers exactly the flagged instances' jobs; the
// key reset runs masked by b.active.
static hipError_t sweep_subset(ps_plan *pl, const Bases &b, ps_map *map_dev, hipStream_t s, ps_plan *pp) {
  std::vector<char> on(pl->n_inst);
  for (int i = 0; i < pl->n_inst; ++i) on[i] = pl->flags_host[i] != 0u;
  struct Seg {
    size_t start = 0;
    int count = 0;
  };
  std::vector<MsgJob> cj;
  auto gather = [&](const Level &L, const WidthClass &c) {
    Seg g;
    g.start = cj.size();
    for (size_t r = 0; r < c.run_inst.size(); ++r)
      if (on[c.run_inst[r]])
        for (int j = 0; j < c.run_count[r]; ++j) cj.push_back(pl->h_jobs[L.off + c.run_first[r] + j]);
    g.count = (int)(cj.size() - g.start);
    return g;
  };
  struct LevelSegs {
    std::vector<Seg> row, col;
    Seg od_row, od_col;
  };
  auto segs_of = [&](const std::vector<Level> &lv) {
    std::vector<LevelSegs> out;
    for (const Level &L : lv) {
      LevelSegs ls;
      for (const WidthClass &c : L.classes) ls.row.push_back(gather(L, c));
      for (const WidthClass &c : L.col_classes) ls.col.push_back(gather(L, c));
      ls.od_row = gather(L, L.od);
      ls.od_col = gather(L, L.od);
      out.push_back(ls);
    }
    return out;
  };
  const std::vector<LevelSegs> up = segs_of(pl->up), down = segs_of(pl->down);
  std::vector<RootJob> cr;
  std::vector<MaxJob> cm;
  std::vector<MapJob> cp;
  for (const RootJob &j : pl->h_roots)
    if (on[j.inst]) cr.push_back(j);
  for (const MaxJob &j : pl->h_maxes)
    if (on[j.inst]) cm.push_back(j);
  for (const MapJob &j : pl->h_maps)
    if (on[j.inst]) cp.push_back(j);
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t o_r = align(cj.size() * sizeof(MsgJob)), o_m = o_r + align(cr.size() * sizeof(RootJob));
  const size_t o_p = o_m + align(cm.size() * sizeof(MaxJob)), n_bytes = o_p + cp.size() * sizeof(MapJob);
  if (n_bytes > pl->cbytes) return hipErrorInvalidValue;
  std::memcpy(pl->cstage, cj.data(), cj.size() * sizeof(MsgJob));
  std::memcpy(pl->cstage + o_r, cr.data(), cr.size() * sizeof(RootJob));
  std::memcpy(pl->cstage + o_m, cm.data(), cm.size() * sizeof(MaxJob));
  std::memcpy(pl->cstage + o_p, cp.data(), cp.size() * sizeof(MapJob));
  hipError_t e = hipMemcpyAsync(pl->ctables, pl->cstage, n_bytes, hipMemcpyHostToDevice, s);
  const MsgJob *jobs = reinterpret_cast<const MsgJob *>(pl->ctables);
  long ev = profile_begin(pp, 0, s);
  if (e == hipSuccess) e = launch_reset(b, pl->n_keys, pl->n_map, pl->n_inst, nullptr, 0, s);
  profile_end(pp, ev, s);
  if (e == hipSuccess) {
    ev = profile_begin(pp, 1, s);
    e = launch_max(b, reinterpret_cast<const MaxJob *>(pl->ctables + o_m), (int)cm.size(), pl->R, pl->max_hw, s);
    profile_end(pp, ev, s);
  }
  auto level = [&](const Level &L, const LevelSegs &ls) {
    for (size_t k = 0; k < L.classes.size() && e == hipSuccess; ++k) {
      const WidthClass &c = L.classes[k];
      const long ev2 = profile_begin(pp, 2, s);
      e = launch_rowpass(b, jobs + ls.row[k].start, ls.row[k].count, pl->taps, pl->shifts, pl->R, c.max_h, c.max_w,
                         L.lds, s, c.row_threads, c.split);
      profile_end(pp, ev2, s);
    }
    if (e == hipSuccess && ls.od_row.count > 0) {
      const long ev2 = profile_begin(pp, 2, s);
      e = launch_rowpass(b, jobs + ls.od_row.start, ls.od_row.count, pl->taps, pl->shifts, pl->R, L.max_h, L.max_w,
                         L.lds, s, 0, L.split);
      profile_end(pp, ev2, s);
    }
    for (size_t k = 0; k < L.col_classes.size() && e == hipSuccess; ++k) {
      const WidthClass &c = L.col_classes[k];
      const long ev2 = profile_begin(pp, 3, s);
      e = launch_colpass(b, jobs + ls.col[k].start, ls.col[k].count, pl->taps, pl->shifts, pl->R, c.epi_hi, c.max_w,
                         L.col_lds, s, c.col_rows, c.split);
      profile_end(pp, ev2, s);
    }
    if (e == hipSuccess && ls.od_col.count > 0) {
      const long ev2 = profile_begin(pp, 3, s);
      e = launch_offdiag(b, jobs + ls.od_col.start, ls.od_col.count, pl->taps, pl->shifts, pl->R, L.od_cells,
                         L.od_hw, s);
      profile_end(pp, ev2, s);
    }
  };
  for (size_t l = 0; l < pl->up.size() && e == hipSuccess; ++l) level(pl->up[l], up[l]);
  if (e == hipSuccess) {
    ev = profile_begin(pp, 4, s);
    e = launch_root(b, reinterpret_cast<const RootJob *>(pl->ctables + o_r), (int)cr.size(), pl->root_children,
                    pl->R, pl->max_hw, false, s);
    profile_end(pp, ev, s);
  }
  for (size_t l = 0; l < pl->down.size() && e == hipSuccess; ++l) level(pl->down[l], down[l]);
  if (e == hipSuccess && map_dev) {
    ev = profile_begin(pp, 5, s);
    e = launch_map_decode(b, reinterpret_cast<const MapJob *>(pl->ctables + o_p), (int)cp.size(), map_dev, s);
    profile_end(pp, ev, s);
  }
  return e;
}

// The plan's method; with the MAP guard on and a BF16 method, every
// instance whose MAP certificate fails (some part's best log marginal beats
// its second best by at most guard_tau) is swept again in FP32, on the
// device, in the same stream order.
static int enqueue(ps_plan *pl, const float *log_unary, float *log_marginal, float *root_posterior, ps_map *map_dev,
                   hipStream_t s, bool profile, const unsigned char *keep = nullptr,
                   const unsigned *fallback = nullptr, bool capturing = false) {
  if (keep) {
    // a pruned execution: the instances without the fallback flag run on
    // their kept tiles, then the flagged ones run the whole sweep
    Bases b;
    b.unary = log_unary;
    b.marginal = log_marginal;
    b.scratch = pl->scratch;
    b.root = root_posterior;
    b.max_keys = pl->keys;
    b.map_keys = pl->map_keys;
    b.top2_keys = nullptr;
    b.active = nullptr;
    b.keep = nullptr;
    ps_plan *pp = profile ? pl : nullptr;
    hipError_t e = hipSuccess;
    if (fallback) {
      e = launch_invert_flags(fallback, pl->not_fallback, pl->n_inst, s);
      b.active = pl->not_fallback;
    }
    b.keep = keep;
    if (e == hipSuccess) e = sweep(pl, b, pl->terms, pl->col_terms, map_dev, s, pp, nullptr);
    if (e == hipSuccess && fallback) {
      b.keep = nullptr;
      b.active = fallback;
      e = sweep(pl, b, pl->terms, pl->col_terms, map_dev, s, pp, nullptr);
    }
    if (e != hipSuccess) return hip_error(e, "kernel launch");
    return PS_OK;
  }
  const bool guard = pl->guard_tau > 0.0f && pl->terms != 0;
  Bases b;
  b.unary = log_unary;
  b.marginal = log_marginal;
  b.scratch = pl->scratch;
  b.root = root_posterior;
  b.max_keys = pl->keys;
  b.map_keys = pl->map_keys;
  b.top2_keys = guard ? pl->top2 : nullptr;
  b.active = nullptr;
  b.keep = nullptr;
  ps_plan *pp = profile ? pl : nullptr;
  hipError_t e = sweep(pl, b, pl->terms, pl->col_terms, map_dev, s, pp, pl->flags);
  if (e == hipSuccess && guard) {
    e = launch_certify(pl->top2, pl->n_map, pl->n_inst, pl->guard_tau, pl->flags, s);
    Bases bf = b;
    bf.top2_keys = nullptr;
    bf.active = pl->flags;
    if (e == hipSuccess && capturing) {
      // under capture the flags stay on the device: every FP32 launch covers
      // the whole batch and the blocks of unflagged instances exit at once
      e = sweep(pl, bf, 0, 0, map_dev, s, pp, nullptr);
    } else if (e == hipSuccess) {
      // eager: read the flags, then launch the FP32 passes over the flagged
      // instances' jobs only
      e = hipMemcpyAsync(pl->flags_host, pl->flags, sizeof(unsigned) * pl->n_inst, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      bool any = false;
      for (int i = 0; i < pl->n_inst && e == hipSuccess; ++i) any = any || pl->flags_host[i] != 0u;
      if (e == hipSuccess && any) e = sweep_subset(pl, bf, map_dev, s, pp);
    }
  }
  if (e != hipSuccess) return hip_error(e, "kernel launch");
  return PS_OK;
}

static int check_exec_args(ps_plan *plan, const float *log_unary, float *log_marginal) {
  if (!plan) return set_error(PS_ERR_INVALID_ARG, "plan is NULL");
  if (!log_unary || !log_marginal) return set_error(PS_ERR_INVALID_ARG, "log_unary and log_marginal are required");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != plan->device)
    return set_error(PS_ERR_INVALID_ARG, "the plan belongs to another device");
  return PS_OK;
}

extern "C" int ps_plan_execute(ps_plan *plan, const float *log_unary, float *log_marginal, float *root_posterior,
                               ps_map *map_dev, void *stream) {
  int st = check_exec_args(plan, log_unary, log_marginal);
  if (st != PS_OK) return st;
  return enqueue(plan, log_unary, log_marginal, root_posterior, map_dev, static_cast<hipStream_t>(stream),
                 plan->profile_mask != 0u);
}

extern "C" int ps_plan_keep_layout(const ps_plan *plan, long long *instance_offsets, long long *total_bytes) {
  if (!plan || !total_bytes) return set_error(PS_ERR_INVALID_ARG, "null argument");
  *total_bytes = plan->keep_bytes;
  if (instance_offsets)
    for (int i = 0; i < plan->n_inst; ++i) instance_offsets[i] = plan->keep_offsets[i];
  return PS_OK;
}

extern "C" int ps_plan_execute_pruned(ps_plan *plan, const float *log_unary, float *log_marginal,
                                      float *root_posterior, ps_map *map_dev, const unsigned char *keep_dev,
                                      const unsigned *fallback_dev, void *stream) {
  int st = check_exec_args(plan, log_unary, log_marginal);
  if (st != PS_OK) return st;
  if (!keep_dev) return set_error(PS_ERR_INVALID_ARG, "keep_dev is required");
  return enqueue(plan, log_unary, log_marginal, root_posterior, map_dev, static_cast<hipStream_t>(stream),
                 plan->profile_mask != 0u, keep_dev, fallback_dev);
}

extern "C" int ps_plan_execute_graph(ps_plan *plan, const float *log_unary, float *log_marginal,
                                     float *root_posterior, ps_map *map_dev, void *stream) {
  int st = check_exec_args(plan, log_unary, log_marginal);
  if (st != PS_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s) return set_error(PS_ERR_INVALID_ARG, "graph execution needs a non-default stream");
  const void *ptrs[4] = {log_unary, log_marginal, root_posterior, map_dev};
  bool same = plan->gexec && plan->gstream == s;
  for (int i = 0; i < 4 && same; ++i) same = plan->gptr[i] == ptrs[i];
  if (!same) {
    if (plan->gexec) {
      (void)hipGraphExecDestroy(plan->gexec);
      plan->gexec = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_error(e, "hipStreamBeginCapture");
    st = enqueue(plan, log_unary, log_marginal, root_posterior, map_dev, s, false, nullptr, nullptr, true);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (st != PS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      plan->gexec = nullptr;
      return hip_error(e, "hipGraphInstantiate");
    }
    plan->gstream = s;
    for (int i = 0; i < 4; ++i) plan->gptr[i] = ptrs[i];
  }
  hipError_t e = hipGraphLaunch(plan->gexec, s);
  if (e != hipSuccess) return hip_error(e, "hipGraphLaunch");
  return PS_OK;
}

// ----------------------------------------------------- reference-shaped API

extern "C" int ps_marginals(const ps_model *model, const ps_instance *instances, int n_instances,
                            const float *log_unary, float *log_marginal, float *root_posterior, ps_map *map_host,
                            void *stream) {
  ps_plan *pl = nullptr;
  int st = ps_plan_create(model, instances, n_instances, &pl);
  if (st != PS_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  ps_map *map_dev = nullptr;
  if (map_host) {
    hipError_t e = hipMalloc(&map_dev, sizeof(ps_map) * pl->n_map);
    if (e != hipSuccess) {
      ps_plan_destroy(pl);
      return set_error(PS_ERR_OUT_OF_MEMORY, "map buffer");
    }
  }
  st = ps_plan_execute(pl, log_unary, log_marginal, root_posterior, map_dev, stream);
  if (st == PS_OK && map_host) {
    hipError_t e = hipMemcpyAsync(map_host, map_dev, sizeof(ps_map) * pl->n_map, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) st = hip_error(e, "map copy");
  }
  if (st == PS_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = hip_error(e, "hipStreamSynchronize");
  }
  if (map_dev) (void)hipFree(map_dev);
  ps_plan_destroy(pl);
  return st;
}

namespace ps {

// One message (computeRotJointMarginal) laid out in one block of device
// memory: its job and max-job records, taps, shifts, the two keys, then the
// row-pass output grid and, for an off-diagonal covariance, the eigenbasis
// grids.
int message_prepare(const ps_model *model, int height, int width, const double offset_c[2], const double offset_p[2],
                    const double C[4], double rot_mean, double rot_sigma, double scale, int is_sparse,
                    MessageSetup &ms) {
  if (!model || !offset_c || !offset_p || !C) return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (height < 1 || width < 1) return set_error(PS_ERR_INVALID_ARG, "sizes must be >= 1");
  const Model &m = model->m;
  const int R = (int)m.ep.num_rotation_steps;
  int device = 0;
  int st = check_device(device);
  if (st != PS_OK) return st;
  MessagePlan mp;
  st = build_message_plan(m.ep, height, width, offset_c, offset_p, C, rot_mean, rot_sigma, scale, is_sparse != 0, mp);
  if (st != PS_OK) return st;
  Builder B(m);
  MsgJob mj;
  std::memset(&mj, 0, sizeof(mj));
  mj.H = height;
  mj.W = width;
  B.add_message(mp, mj);
  ms.R = R;
  ms.H = height;
  ms.W = width;
  ms.od = mp.od != 0;
  ms.split = mp.split;
  ms.od_cells = mp.od ? (long long)R * mp.th * mp.tw : 0;
  ms.lds = rowpass_lds_bytes(R, width, mj.kx);
  ms.col_lds = mp.od ? 0 : colpass_lds_bytes(mp.ky, 16);
  if (ms.lds > kLdsLimit || ms.col_lds > kLdsLimit)
    return set_error(PS_ERR_UNSUPPORTED, "message needs more LDS than one workgroup has");
  mj.src_kind = kSrcUnary;  // a plain grid, read through the (empty) mask path
  mj.src = gref(kBufUnary, 0, width);
  mj.src_key = 0;
  mj.tmp = gref(kBufScratch, 0, t_pitch(width));
  mj.dst_kind = kDstUpRoot;
  mj.dst = gref(kBufMarginal, 0, width);
  mj.dst_unary = kNone;
  mj.dst2 = kNone;
  mj.dst_bel = kNone;
  const long long n64 = ((long long)R * height * t_pitch(width) + 63) & ~63ll;
  mj.od_buf = mp.od ? gref(kBufScratch, n64, 0) : kNone;
  ms.epi = epi_kind(mj);
  MaxJob xj;
  std::memset(&xj, 0, sizeof(xj));
  xj.H = height;
  xj.W = width;
  xj.src_kind = kSrcUnary;
  xj.key = 0;
  xj.src = gref(kBufUnary, 0, width);

  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  ms.o_job = 0;
  ms.o_max = align(sizeof(MsgJob));
  ms.o_taps = align(ms.o_max + sizeof(MaxJob));
  ms.o_sh = align(ms.o_taps + B.taps.size() * sizeof(float));
  ms.o_key = align(ms.o_sh + B.shifts.size() * sizeof(int));
  ms.o_mk = align(ms.o_key + sizeof(unsigned));
  ms.o_T = align(ms.o_mk + sizeof(unsigned long long));
  ms.total = ms.o_T + sizeof(float) * (size_t)(n64 + (mp.od ? od_scratch_floats(R, mp.th, mp.tw) : 0));
  ms.host.assign(ms.o_T, 0);
  std::memcpy(ms.host.data() + ms.o_job, &mj, sizeof(mj));
  std::memcpy(ms.host.data() + ms.o_max, &xj, sizeof(xj));
  std::memcpy(ms.host.data() + ms.o_taps, B.taps.data(), B.taps.size() * sizeof(float));
  std::memcpy(ms.host.data() + ms.o_sh, B.shifts.data(), B.shifts.size() * sizeof(int));
  return PS_OK;
}

// Upload the tables into `mem` (at least ms.total bytes of device memory),
// run the message and wait for it.  The two keys are cleared by a kernel.
int message_enqueue(const MessageSetup &ms, char *mem, const float *child_dev, float *parent_dev, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(mem, ms.host.data(), ms.o_T, hipMemcpyHostToDevice, s);
  Bases b;
  b.unary = child_dev;
  b.marginal = parent_dev;
  b.scratch = reinterpret_cast<float *>(mem + ms.o_T);
  b.root = nullptr;
  b.max_keys = reinterpret_cast<unsigned *>(mem + ms.o_key);
  b.map_keys = reinterpret_cast<unsigned long long *>(mem + ms.o_mk);
  b.top2_keys = nullptr;
  b.active = nullptr;
  b.keep = nullptr;
  const MsgJob *jd = reinterpret_cast<const MsgJob *>(mem + ms.o_job);
  const float *td = reinterpret_cast<const float *>(mem + ms.o_taps);
  const int *sd = reinterpret_cast<const int *>(mem + ms.o_sh);
  if (e == hipSuccess) e = rowpass_set_lds_limit(ms.lds);
  if (e == hipSuccess && !ms.od) e = colpass_set_lds_limit(ms.col_lds);
  if (e == hipSuccess) e = launch_reset(b, 1, 1, 1, nullptr, 0, s);
  if (e == hipSuccess)
    e = launch_max(b, reinterpret_cast<const MaxJob *>(mem + ms.o_max), 1, ms.R, ms.H * ms.W, s);
  if (e == hipSuccess) e = launch_rowpass(b, jd, 1, td, sd, ms.R, ms.H, ms.W, ms.lds, s, 0, ms.split);
  if (e == hipSuccess) {
    if (ms.od)
      e = launch_offdiag(b, jd, 1, td, sd, ms.R, ms.od_cells, (long long)ms.H * ms.W, s);
    else
      e = launch_colpass(b, jd, 1, td, sd, ms.R, ms.epi, ms.W, ms.col_lds, s, 0, ms.split);
  }
  // the caller's tables live in pageable host memory and the block may be
  // reused right after: the call is synchronous
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_error(e, "message");
  return PS_OK;
}

}  // namespace ps

extern "C" int ps_message(const ps_model *model, const float *child_dev, float *parent_dev, int height, int width,
                          const double offset_c[2], const double offset_p[2], const double C[4], double rot_mean,
                          double rot_sigma, double scale, int is_sparse, void *stream) {
  if (!child_dev || !parent_dev) return set_error(PS_ERR_INVALID_ARG, "null argument");
  MessageSetup ms;
  int st = message_prepare(model, height, width, offset_c, offset_p, C, rot_mean, rot_sigma, scale, is_sparse, ms);
  if (st != PS_OK) return st;
  void *mem = nullptr;
  if (hipMalloc(&mem, ms.total) != hipSuccess) return set_error(PS_ERR_OUT_OF_MEMORY, "message scratch");
  st = message_enqueue(ms, static_cast<char *>(mem), child_dev, parent_dev, static_cast<hipStream_t>(stream));
  (void)hipFree(mem);
  return st;
}

// --------------------------------------------------------------- utilities

extern "C" const char *ps_last_error_string(void) { return g_last_error.c_str(); }

extern "C" const char *ps_version(void) { return "partapp_ps 0.1 (gfx950)"; }

extern "C" int ps_debug_message_plan(const ps_exp_param *ep, int height, int width, const double offset_c[2],
                                     const double offset_p[2], const double C[4], double rot_mean,
                                     double rot_sigma, double scale,
                                     int *delta, int *fwd_ox, int *fwd_oy, int *inv_ox, int *inv_oy, int *rot_len,
                                     float *rot_taps, int *kx, int *ky) {
  if (!ep || !offset_c || !offset_p || !C || !delta || !fwd_ox || !fwd_oy || !inv_ox || !inv_oy || !rot_len || !kx ||
      !ky)
    return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (ep->num_rotation_steps < 1 || ep->num_rotation_steps > (unsigned)kMaxRot)
    return set_error(PS_ERR_UNSUPPORTED, "num_rotation_steps must be in [1, 64]");
  MessagePlan mp;
  int st = build_message_plan(*ep, height, width, offset_c, offset_p, C, rot_mean, rot_sigma, scale, false, mp);
  if (st != PS_OK) return st;
  const int R = (int)ep->num_rotation_steps;
  *delta = mp.delta;
  for (int r = 0; r < R; ++r) {
    fwd_ox[r] = mp.fwd_ox[r];
    fwd_oy[r] = mp.fwd_oy[r];
    inv_ox[r] = mp.inv_ox[r];
    inv_oy[r] = mp.inv_oy[r];
  }
  *rot_len = mp.rot_mode == kRotFilter ? (int)mp.rot_taps.size() : 0;
  if (rot_taps)
    for (size_t i = 0; i < mp.rot_taps.size(); ++i) rot_taps[i] = mp.rot_taps[i];
  *kx = mp.kx;
  *ky = mp.ky;
  return PS_OK;
}
