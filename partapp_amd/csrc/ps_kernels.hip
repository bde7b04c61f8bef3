// gfx950 kernels of the pictorial-structure sweep.
//
// One message of the sum-product sweep (computeRotJointMarginal,
// objectdetect_findrot.cpp:292-456) is two launches over a whole level of
// messages (every instance, scale and chain of the level at once):
//
//   k_rowpass  one block per (message, row y): gathers the forward-warped
//              child belief of all R rotation slices (integer shift per
//              slice), exp(b - M), circular rotation filter, then the
//              truncated x filter from an LDS row with zero halo.  Writes the
//              row-filtered grid T.
//   k_colpass  one block per (message, rotation, 64x64 tile): truncated y
//              filter of T with a 16-row register window per lane, log,
//              +M, the inverse warp folded into the tile origin, and the
//              caller's accumulation (parent belief, root message slot, or
//              marginal + next downward source) fused as the epilogue,
//              together with the max / argmax reductions the next step needs.
//
// Arithmetic follows the reference order: every dot product accumulates its
// taps in ascending order with a fused multiply-add starting from 0, exactly
// like a sequential sdot; sums of grids are added in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>

#include "ps_device.hpp"
#include "ps_rowgather.hpp"
#include "ps_internal.hpp"
#include "../../include/partapp_ps.h"

namespace ps {

// --------------------------------------------------------------- k_reset

// Clears the max keys, the argmax keys and their top-2 keys of the instances
// that run (keys_per_inst / map_per_inst consecutive entries per instance);
// with clear_active it also clears the per-instance flags (n_flags) that a
// later certification pass sets.
__global__ void k_reset(Bases b, int n_keys, int n_map, int keys_per_inst, int map_per_inst, unsigned *flags,
                        int n_flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_keys && !skip_inst(b, i / keys_per_inst)) b.max_keys[i] = 0u;
  if (i < n_map && !skip_inst(b, i / map_per_inst)) {
    b.map_keys[i] = 0ull;
    if (b.top2_keys) b.top2_keys[i] = 0ull;
  }
  if (flags && i < n_flags) flags[i] = 0u;
}

// ------------------------------------------------------------------ k_max
// Max of a leaf's source grid (getMinMax, multi_array_op.hpp:60-77).  Grids
// without masks are scanned flat with 16-byte loads; masked ones by rows.

__global__ __launch_bounds__(256) void k_max(Bases b, const MaxJob *__restrict__ jobs, int R, int chunk) {
  const MaxJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  const int H = J.H, W = J.W;
  const long long n = (long long)R * H * W;
  const long long e0 = (long long)blockIdx.x * chunk;
  unsigned km = 0u;
  if (e0 < n) {
    const long long e1 = min(e0 + (long long)chunk, n);
    if (J.src_kind == kSrcZero) {
      km = f2key(0.0f);
    } else {
      const float *src = cptr(b, J.src);
      const bool plain = J.mask.rot_bits == 0ull && J.mask.strip == 0;
      // four independent loads per thread in flight (a lane past the end
      // reads its first element again, which leaves the max unchanged)
      constexpr int kU = 4;
      if (plain && ((reinterpret_cast<uintptr_t>(src + e0) & 15) == 0) && ((e1 - e0) & 3) == 0) {
        const float4 *s4 = reinterpret_cast<const float4 *>(src + e0);
        const int n4 = (int)((e1 - e0) >> 2);
        for (int i = threadIdx.x; i < n4; i += kU * blockDim.x) {
          float4 v[kU];
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            const int j = i + u * (int)blockDim.x;
            v[u] = s4[j < n4 ? j : i];
          }
#pragma unroll
          for (int u = 0; u < kU; ++u)
            km = max(km, max(max(f2key(v[u].x), f2key(v[u].y)), max(f2key(v[u].z), f2key(v[u].w))));
        }
      } else {
        const long long HW = (long long)H * W;
        for (long long e = e0 + threadIdx.x; e < e1; e += kU * blockDim.x) {
          float v[kU];
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            const long long f = e + u * (long long)blockDim.x;
            const long long g = f < e1 ? f : e;
            const int r = (int)(g / HW);
            const int x = (int)(g % W);
            v[u] = masked(J.mask, r, x, W) ? kLogZero : src[g];
          }
#pragma unroll
          for (int u = 0; u < kU; ++u) km = max(km, f2key(v[u]));
        }
      }
    }
  }
  km = wave_max_u32(km);
  if ((threadIdx.x & 63) == 0 && km != 0u) atomicMax(b.max_keys + J.key, km);
}

// -------------------------------------------------------------- k_rowpass
// One block per (message, row y), all R slices of the row.  LDS holds
// P[R][Pw]: the rotation-filtered rows with kx zeros on the left and a zero
// tail, Pw a multiple of 4, read by the x filter.

// slice groups the fixed-R row pass stages in LDS one after another: 48
// slices in two halves (36 KB instead of 72 KB of LDS per 300-wide row, 160
// VGPRs: 20.9 instead of 27.6 ms per 32 x 300x300x48 sweep), 24 in one (as
// two halves 9.7 instead of 8.8 ms per 32 x 300x300x24 sweep)
__host__ __device__ constexpr int row_groups(int R) { return R >= 48 ? 2 : 1; }

struct RowGeom {
  int Wo;    // groups of 8 outputs per row
  int nch;   // chunks of 8 taps (the last one a half chunk when tail4)
  int Pw;    // padded row pitch of P, a multiple of 8
  bool tail4;  // the taps, padded to a multiple of 4, end in a chunk of 4
};
__host__ __device__ __forceinline__ RowGeom row_geom(int W, int kx) {
  RowGeom g;
  g.Wo = (W + 7) >> 3;
  g.nch = (2 * kx + 1 + 7) >> 3;
  g.Pw = 8 * (g.Wo + g.nch + 1);
  g.tail4 = ((2 * kx + 1 + 3) & 7) == 4;
  return g;
}

// zero the halo columns [0, kx) and [kx + W, Pw) of every row of P
__device__ __forceinline__ void zero_halo(float *P, int R, int W, int kx, const RowGeom &g) {
  const int per_row = g.Pw - W;
  for (int i = threadIdx.x; i < R * per_row; i += blockDim.x) {
    const int r = i / per_row;
    const int c = i - r * per_row;
    P[r * g.Pw + (c < kx ? c : c + W)] = 0.0f;
  }
}

// truncated x filter (multi_array_filter.hpp:277-288) of every row of P:
// eight outputs per lane, taps in chunks of eight (zero padded), a 16-float
// window sliding along the row in two 16-byte LDS reads per chunk; writes
// row y of T (row pitch tp, a multiple of 32: the outputs past W land in the
// padding)
// NCH chunks known at compile time: the chunk loop unrolls, so the sliding
// window is renamed instead of copied.  With g.tail4 the last chunk runs its
// first 4 taps only (the others are the zero padding: an fmaf with a zero tap
// leaves the sum as it is)
template <int NCH>
__device__ __forceinline__ void xfilter_rows_n(const float *P, const float *taps_x, float *T, int R, int tp,
                                               long long plane, int y, const RowGeom &g) {
  const bool tail4 = g.tail4;
  const int n_items = R * g.Wo;
  const float inv_wo = 1.0f / (float)g.Wo;
  for (int item = threadIdx.x; item < n_items; item += blockDim.x) {
    int ro = (int)((float)item * inv_wo);
    if (ro * g.Wo > item) --ro;
    if ((ro + 1) * g.Wo <= item) ++ro;
    const int q = item - ro * g.Wo;
    // rows of P start on 32-byte boundaries (Pw is a multiple of 8)
    const float4 *P4 = reinterpret_cast<const float4 *>(
        __builtin_assume_aligned(P + ro * g.Pw + 8 * q, 16));
    float a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = 0.0f;
    float4 w0 = P4[0], w1 = P4[1];
    // the taps are re-read per item through the scalar cache rather than held
    // in vector registers for the whole loop
    const float *tx = taps_x;
    asm volatile("" : "+s"(tx));
#pragma unroll
    for (int c = 0; c < NCH - 1; ++c) {
      const float4 w2 = P4[2 * c + 2], w3 = P4[2 * c + 3];
      const float win[16] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                             w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
      float f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = tx[8 * c + u];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = fmaf(win[i + u], f[u], a[i]);
      w0 = w2;
      w1 = w3;
    }
    {  // the last chunk: 4 or 8 taps
      constexpr int c = NCH - 1;
      const float4 w2 = P4[2 * c + 2];
      float f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = tx[8 * c + u];
      if (tail4) {
        const float win[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = fmaf(win[i + u], f[u], a[i]);
      } else {
        const float4 w3 = P4[2 * c + 3];
        const float win[16] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                               w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = fmaf(win[i + u], f[u], a[i]);
      }
    }
    float *trow = T + ro * plane + (long long)y * tp;
    const int x0 = 8 * q;
    *reinterpret_cast<float4 *>(trow + x0) = make_float4(a[0], a[1], a[2], a[3]);
    *reinterpret_cast<float4 *>(trow + x0 + 4) = make_float4(a[4], a[5], a[6], a[7]);
  }
}

__device__ __forceinline__ void xfilter_rows(const float *P, const float *taps_x, float *T, int R, int tp,
                                             long long plane, int y, const RowGeom &g) {
  switch (g.nch) {  // half widths up to 43 (the filters of sigma <= 14.3 px)
    case 1: xfilter_rows_n<1>(P, taps_x, T, R, tp, plane, y, g); return;
    case 2: xfilter_rows_n<2>(P, taps_x, T, R, tp, plane, y, g); return;
    case 3: xfilter_rows_n<3>(P, taps_x, T, R, tp, plane, y, g); return;
    case 4: xfilter_rows_n<4>(P, taps_x, T, R, tp, plane, y, g); return;
    case 5: xfilter_rows_n<5>(P, taps_x, T, R, tp, plane, y, g); return;
    case 6: xfilter_rows_n<6>(P, taps_x, T, R, tp, plane, y, g); return;
    case 7: xfilter_rows_n<7>(P, taps_x, T, R, tp, plane, y, g); return;
    case 8: xfilter_rows_n<8>(P, taps_x, T, R, tp, plane, y, g); return;
    case 9: xfilter_rows_n<9>(P, taps_x, T, R, tp, plane, y, g); return;
    case 10: xfilter_rows_n<10>(P, taps_x, T, R, tp, plane, y, g); return;
    case 11: xfilter_rows_n<11>(P, taps_x, T, R, tp, plane, y, g); return;
    default: break;
  }
  const int n_items = R * g.Wo;
  const float inv_wo = 1.0f / (float)g.Wo;
  for (int item = threadIdx.x; item < n_items; item += blockDim.x) {
    int ro = (int)((float)item * inv_wo);
    if (ro * g.Wo > item) --ro;
    if ((ro + 1) * g.Wo <= item) ++ro;
    const int q = item - ro * g.Wo;
    const float4 *P4 = reinterpret_cast<const float4 *>(P + ro * g.Pw) + 2 * q;
    float a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = 0.0f;
    float4 w0 = P4[0], w1 = P4[1];
    for (int c = 0; c < g.nch; ++c) {
      const float4 w2 = P4[2 * c + 2], w3 = P4[2 * c + 3];
      const float win[16] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                             w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
      float f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = taps_x[8 * c + u];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = fmaf(win[i + u], f[u], a[i]);
      w0 = w2;
      w1 = w3;
    }
    float *trow = T + ro * plane + (long long)y * tp;
    const int x0 = 8 * q;
    *reinterpret_cast<float4 *>(trow + x0) = make_float4(a[0], a[1], a[2], a[3]);
    *reinterpret_cast<float4 *>(trow + x0 + 4) = make_float4(a[4], a[5], a[6], a[7]);
  }
}

// Forward offsets of output slice ro (so = J.shifts + 4 R: the x offsets at
// so, the y offsets at so + R) at row ys / column x.  SPLIT: the warp's
// offset changes once along an axis, by the step at so + 5 R (x) / so + 7 R
// (y) from the break at so + 4 R / so + 6 R on.
template <int R, bool SPLIT>
__device__ __forceinline__ int fwd_oy(const int *__restrict__ shifts, int so, int ro, int ys) {
  int oy = shifts[so + R + ro];
  if (SPLIT && ys >= shifts[so + 6 * R + ro]) oy += shifts[so + 7 * R + ro];
  return oy;
}
template <int R, bool SPLIT>
__device__ __forceinline__ int fwd_ox(const int *__restrict__ shifts, int so, int ro, int x) {
  int ox = shifts[so + ro];
  if (SPLIT && x >= shifts[so + 4 * R + ro]) ox += shifts[so + 5 * R + ro];
  return ox;
}

// One column x of the fixed-R row pass for slice group G: the R source
// values are loaded together, exponentiated and rotation-filtered in
// registers; the group's output slices G * RG .. (G + 1) * RG - 1 are written
// to P.
template <int R, int G, bool SPLIT>
__device__ __forceinline__ void rowpass_column(const MsgJob &J, const float *__restrict__ taps,
                                               const int *__restrict__ shifts, const float *src, float negM,
                                               float e_fill, float *P, int y, const RowGeom &g, const int x) {
  constexpr int RG = R / row_groups(R);
  const int H = J.H, W = J.W;
  const int kx = J.kx;
  const float *fr = taps + J.taps_r;
  const int delta = J.delta;
  const bool zero_src = J.src_kind == kSrcZero;
  const UnaryMask mk = J.src_mask;
  const bool use_mask = J.src_kind == kSrcUnary && (mk.rot_bits != 0ull || mk.strip != 0);
  const int Ps = J.src.pitch;
  int HW = H * Ps;  // R * H * pitch < 2^31 is checked when the plan is built
  // per OUTPUT slice: the source slice is ro - delta (clamped into [0, R) at
  // so + 3 R), read at (y + oyo, x + oxo) with oxo at so, oyo at so + R;
  // slices without a source carry a row offset that fails the bounds test.
  // The offsets are read through the scalar cache twice (an empty asm keeps
  // the reads apart) rather than held in scalar registers from the loads to
  // their use.
  int so = J.shifts + 4 * R;
  const int ys = __builtin_amdgcn_readfirstlane(y);  // uniform: the row offsets on the scalar unit

  // forward warp (objectdetect_findrot.cpp:342-359): every load unconditional
  // (a slice without a source reads its clamped slice, a column outside the
  // grid its nearest column), all R of them issued before any is used; the
  // row address on the scalar unit, the column a byte offset from it
  float v[R];
  asm volatile("" : "+s"(so), "+s"(HW));
  const int x4 = x * 4, xmax4 = (W - 1) * 4;
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    const int yy = min(max(ys + fwd_oy<R, SPLIT>(shifts, so, ro, ys), 0), H - 1);
    const char *srow = reinterpret_cast<const char *>(src) + (unsigned)(shifts[so + 3 * R + ro] * HW + yy * Ps) * 4u;
    const int xx4 = min(max(x4 + fwd_ox<R, SPLIT>(shifts, so, ro, x) * 4, 0), xmax4);
    v[ro] = *reinterpret_cast<const float *>(srow + (unsigned)xx4);
  }
  // masks, then exp(b - M) (:362-365) as 2^((b - M) log2 e) on the hardware
  // exp2 (values below 2^-126 flush to zero, where the reference's expf
  // returns subnormals).  A source pixel outside the grid (the warp's
  // LOG_ZERO fill) and a masked unary state (a masked rotation, or a column
  // in the masked strips) both take e_fill = exp(LOG_ZERO - M): a masked
  // rotation fails its row's test, the strips narrow the columns' range.
  asm volatile("" : "+s"(so));
  const int lo = use_mask ? mk.strip : 0;
  const unsigned span = (unsigned)max(W - 2 * lo, 0);  // columns lo .. W - lo - 1 are read
  const unsigned long long rot_bits = use_mask ? mk.rot_bits : 0ull;
  float e[R];
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    const int rs = ro - delta;
    const bool row_ok = (unsigned)(ys + fwd_oy<R, SPLIT>(shifts, so, ro, ys)) < (unsigned)H &&
                        !((unsigned)rs < (unsigned)R && ((rot_bits >> (rs & 63)) & 1ull));
    // a leaf without a detector has the zero log posterior (b = 0); the
    // exponential is computed for every lane (an empty asm keeps it out of a
    // branch), then selected
    float ev = __builtin_amdgcn_exp2f(((zero_src ? 0.0f : v[ro]) + negM) * kLog2eF);
    asm volatile("" : "+v"(ev));
    const bool ok = row_ok && (unsigned)(x + fwd_ox<R, SPLIT>(shifts, so, ro, x) - lo) < span;
    e[ro] = ok ? ev : e_fill;
  }
  float *out = P + kx + x;
  if (J.rot_mode == kRotFilter) {
    RotDispatch<R, (R % 2 == 1) ? R - 2 : R - 1, G * RG, (G + 1) * RG>::run(J.rot_len, e, fr, out, g.Pw);
  } else if (J.rot_mode == kRotCopy) {
#pragma unroll
    for (int k = 0; k < RG; ++k) out[k * g.Pw] = e[G * RG + k];
  } else {
#pragma unroll
    for (int k = 0; k < RG; ++k) out[k * g.Pw] = 0.0f;
  }
}

// Row pass for a rotation count fixed at compile time (24, and the
// reference's default 48): lane t owns the columns x = t (mod blockDim.x);
// the slice groups go through P and the x filter one after another.
template <int R, bool SPLIT>
__global__ __launch_bounds__(512) void k_rowpass_r(Bases b, const MsgJob *__restrict__ jobs,
                                                   const float *__restrict__ taps, const int *__restrict__ shifts) {
  extern __shared__ float4 lds4[];
  float *P = reinterpret_cast<float *>(lds4);
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int y = (int)bx;
  const int H = J.H, W = J.W;
  if (y >= H) return;
  const int kx = J.kx;
  const RowGeom g = row_geom(W, kx);
  constexpr int RG = R / row_groups(R);
  zero_halo(P, RG, W, kx, g);

  const float M = key2f(b.max_keys[J.src_key]);
  const float negM = -M;
  // exp(LOG_ZERO - M): the warp's fill value, on the path of every other value
  const float e_fill = __builtin_amdgcn_exp2f((kLogZero + negM) * kLog2eF);
  const float *src = cptr(b, J.src);
  float *T = mptr(b, J.tmp);
  const long long plane = (long long)H * J.tpitch;
  // a warp whose offset changes along an axis takes the SPLIT variants
  const bool split = SPLIT && J.split != 0;
  for (int x = threadIdx.x; x < W; x += blockDim.x) {
    if (split) rowpass_column<R, 0, true>(J, taps, shifts, src, negM, e_fill, P, y, g, x);
    else rowpass_column<R, 0, false>(J, taps, shifts, src, negM, e_fill, P, y, g, x);
  }
  __syncthreads();
  xfilter_rows(P, taps + J.taps_x, T, RG, J.tpitch, plane, y, g);
  if constexpr (row_groups(R) > 1) {
    __syncthreads();  // every item of the first group is filtered
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
      if (split) rowpass_column<R, 1, true>(J, taps, shifts, src, negM, e_fill, P, y, g, x);
      else rowpass_column<R, 1, false>(J, taps, shifts, src, negM, e_fill, P, y, g, x);
    }
    __syncthreads();
    xfilter_rows(P, taps + J.taps_x, T + RG * plane, RG, J.tpitch, plane, y, g);
  }
}

// Row pass for any rotation count: the exponentiated column goes through an
// LDS buffer E[R][Wa] before the rotation filter.
__global__ __launch_bounds__(512) void k_rowpass_any(Bases b, const MsgJob *__restrict__ jobs,
                                                     const float *__restrict__ taps, const int *__restrict__ shifts,
                                                     int R) {
  extern __shared__ float4 lds4[];
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int y = (int)bx;
  const int H = J.H, W = J.W;
  if (y >= H) return;
  const int kx = J.kx;
  const RowGeom g = row_geom(W, kx);
  const int Wa = (W + 3) & ~3;
  float *E = reinterpret_cast<float *>(lds4);
  float *P = E + R * Wa;
  zero_halo(P, R, W, kx, g);

  const float M = key2f(b.max_keys[J.src_key]);
  const float negM = -M;
  const float e_fill = expf(kLogZero + negM);
  const float *src = cptr(b, J.src);
  const int Ps = J.src.pitch;
  const long long HW = (long long)H * Ps;
  const int *fox = shifts + J.shifts;
  const int *foy = fox + R;
  const int n_rot = J.rot_len >> 1;
  const float *fr = taps + J.taps_r;
  for (int x = threadIdx.x; x < W; x += blockDim.x) {
    for (int ro = 0; ro < R; ++ro) {
      const int rs = ro - J.delta;
      float e = e_fill;
      if (rs >= 0 && rs < R) {
        int oy = foy[rs], ox = fox[rs];
        if (J.split) {  // the breaks and steps by output slice
          const int *sp = shifts + J.shifts + 8 * R;
          if (x >= sp[ro]) ox += sp[R + ro];
          if (y >= sp[2 * R + ro]) oy += sp[3 * R + ro];
        }
        const int yy = y + oy;
        const int xx = x + ox;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
          float v;
          if (J.src_kind == kSrcZero) v = 0.0f;
          else if (J.src_kind == kSrcUnary && masked(J.src_mask, rs, xx, W)) v = kLogZero;
          else v = src[rs * HW + (long long)yy * Ps + xx];
          e = __builtin_amdgcn_exp2f((v + negM) * kLog2eF);
        }
      }
      E[ro * Wa + x] = e;
    }
    // the column was written by this thread: no barrier needed
    for (int ro = 0; ro < R; ++ro) {
      float acc;
      if (J.rot_mode == kRotFilter) {
        acc = 0.0f;
        for (int j = 0; j < J.rot_len; ++j) {
          int rr = ro - n_rot + j;
          rr = rr < 0 ? rr + R : (rr >= R ? rr - R : rr);
          acc = fmaf(E[rr * Wa + x], fr[j], acc);
        }
      } else if (J.rot_mode == kRotCopy) {
        acc = E[ro * Wa + x];
      } else {
        acc = 0.0f;
      }
      P[ro * g.Pw + kx + x] = acc;
    }
  }
  __syncthreads();
  xfilter_rows(P, taps + J.taps_x, mptr(b, J.tmp), R, J.tpitch, (long long)H * J.tpitch, y, g);
}

// -------------------------------------------------------------- k_colpass
// One block per (message, slice r, strip of 64 columns); it walks down the
// strip in tiles of 4 ROWS output rows.  Each tile stages the source rows it
// needs (4 ROWS + padded taps, zero outside the grid, inverse shift folded
// into the origin) in LDS; each of the four waves then computes ROWS rows,
// one column per lane, with a ROWS-row register ring fed from LDS.  Taps run
// in chunks of 8 (the tap pool is zero padded), so the ring indices are
// static.  ROWS = 8 (the default: at most 79 VGPRs, 6 waves per SIMD) or 16
// (twice the filter work per LDS read, at most 128 VGPRs and 4 waves); the
// outputs are the same bit for bit.

constexpr int kColTileX = 64;
constexpr int kColWaves = 4;
constexpr int kColRowsDefault = 8;

__host__ __device__ __forceinline__ int col_padded_taps(int ky) { return (2 * ky + 1 + 3) & ~3; }
__host__ __device__ __forceinline__ int col_tile_rows(int ky, int rows) { return kColWaves * rows + col_padded_taps(ky); }
// rows moved per thread and pass when a tile hands its halo to the next one
constexpr int kColMove = 24;

// N (8 or 4) taps of the y filter starting at tap t; PH is the ring phase (t % ROWS)
template <int PH, int N, int ROWS>
__device__ __forceinline__ void col_chunk(float (&acc)[ROWS], float (&buf)[ROWS], const float *fy, const float *tw,
                                          int t) {
  float f[N];
#pragma unroll
  for (int u = 0; u < N; ++u) f[u] = fy[t + u];
#pragma unroll
  for (int u = 0; u < N; ++u) {
#pragma unroll
    for (int i = 0; i < ROWS; ++i) acc[i] = fmaf(buf[(i + u + PH) & (ROWS - 1)], f[u], acc[i]);
    buf[(u + PH) & (ROWS - 1)] = tw[(t + u + ROWS) * kColTileX];
  }
}

// The caller's accumulation is a template (EPI, one variant per destination
// kind), so every load of a tile is unconditional straight-line code: the
// epilogue operands first, then the next tile's rows (always loaded, clamped
// into the grid, zeroed when they reach LDS), and the epilogue waits for its
// operands only.
// SPLIT: the launch may hold a split warp (J.split), whose inverse offsets
// change once along x (per lane) or y (a second segment of rows)
template <int EPI, int ROWS, bool SPLIT>
__device__ __forceinline__ void colpass_strip(const Bases &b, const MsgJob &J, const float *__restrict__ taps,
                                              const int *__restrict__ shifts, int R, int r, int xt, float *tile) {
  constexpr int kColRows = ROWS;
  constexpr int kColTileY = kColRows * kColWaves;
  constexpr bool kNeedU = Epi<EPI>::kNeedU;
  constexpr bool kNeedB = Epi<EPI>::kNeedB;
  const int H = J.H, W = J.W;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = xt * kColTileX + lane;
  const bool xin = x < W;
  const int xo = xin ? x : W - 1;

  // inverse offsets of slice r (a split warp: the lane's column past the x
  // break takes the x step, the rows from the y break on the y step)
  int ox = shifts[J.shifts + 2 * R + r];
  const int oy0 = shifts[J.shifts + 3 * R + r];
  const bool split = SPLIT && J.split != 0;
  if (split && x >= shifts[J.shifts + 12 * R + r]) ox += shifts[J.shifts + 13 * R + r];
  const int ybrk = split ? shifts[J.shifts + 14 * R + r] : H;
  const int ystep = split ? shifts[J.shifts + 15 * R + r] : 0;
  const bool col_ok = xin && x + ox >= 0 && x + ox < W;
  int tp = J.tpitch;
  const float *Tr = cptr(b, J.tmp) + (long long)r * H * tp;
  const int ky = J.ky;
  const int Lp = col_padded_taps(ky);
  const int nrows = col_tile_rows(ky, ROWS);
  const float *fy = taps + J.taps_y;
  const float M = key2f(b.max_keys[J.src_key]);
  // row pitches: pa of dst (and of the unary a marginal adds), pb of the
  // grids at the other pitch (EpiPitch)
  const int pa = J.dst.pitch, pb = epi_pitch_b<EPI>(J);
  float *dst = mptr(b, J.dst) + r * H * pa;
  const float *dun = kNeedU ? cptr(b, J.dst_unary) + r * H * J.dst_unary.pitch : dst;
  float *dst2 = kNeedB ? mptr(b, J.dst2) + r * H * pb : dst;
  const float *bel = kNeedB ? cptr(b, J.dst_bel) + r * H * pb : dst;
  const bool leaf_u = J.dst_base == 1;  // a leaf's marginal: detected part (else its base is 0)
  const bool xmask = ((J.dst_mask.rot_bits >> r) & 1ull) || x < J.dst_mask.strip || x >= W - J.dst_mask.strip;
  // column loaded by this thread's lane slot (loads outside the grid go to its
  // nearest row and column; those values are zeroed when they reach LDS)
  const int lcol = xt * kColTileX + lane + ox;
  const bool lcol_ok = x < W && lcol >= 0 && lcol < W;
  const float *Tc = Tr + min(max(lcol, 0), W - 1);
  auto src_raw = [&](int row) -> float { return Tc[(long long)min(max(row, 0), H - 1) * tp]; };
  auto src_ok = [&](int row) -> bool { return lcol_ok && row >= 0 && row < H; };

  // The output rows run in segments of one inverse row offset oy: [0, ybrk)
  // and, for a split warp, [ybrk, H).  Within a segment, tile t holds source
  // rows s_t + [0, nrows), s_t = ylo + 64 t + oy - ky.  Tile t+1 shares rows
  // [64, nrows) of tile t (moved to [0, Lp) in LDS) and needs 64 new rows,
  // fetched into registers while tile t is filtered.
  constexpr int kNew = kColTileY / kColWaves;  // new rows per thread
  Epi<EPI> acc;
  // (the padding columns of the plan's grids are left unwritten here: this
  // pass is bound by its FMAs, and the extra stores cost more than the whole
  // lines save)
  unsigned Wb = (unsigned)pa * 4u, Wb2 = (unsigned)pb * 4u;
  for (int seg = 0; seg < (SPLIT ? 2 : 1); ++seg) {
    const int ylo = seg == 0 ? 0 : ybrk, yhi = seg == 0 ? ybrk : H;
    if (ylo >= yhi) continue;
    const int oy = seg == 0 ? oy0 : oy0 + ystep;
    if (seg) __syncthreads();  // every wave is done with the first segment's tile
    for (int rr = wave; rr < nrows; rr += kColWaves) {
      const int row = ylo + oy - ky + rr;
      const float v = src_raw(row);
      tile[rr * kColTileX + lane] = src_ok(row) ? v : 0.0f;
    }
    __syncthreads();

    for (int ty0 = ylo; ty0 < yhi; ty0 += kColTileY) {
      asm volatile("" : "+s"(Wb), "+s"(Wb2), "+s"(tp));  // per-row offsets recomputed per tile, not hoisted
      const bool more = ty0 + kColTileY < yhi;
      const int y0 = ty0 + wave * kColRows;
      const unsigned eb0 = (unsigned)(y0 * pa + xo) * 4u;
      const unsigned ebs0 = Epi<EPI>::kSide ? (unsigned)(y0 * pb + xo) * 4u : eb0;
      // the unary of the part written, loaded now so that its latency hides
      // behind the filter; the stored belief of the five-grid downward epilogue
      // is loaded after the filter (held through it, its 16 registers make the
      // kernel spill at 128 VGPRs)
      float pu[kColRows], pbel[kColRows];
      if (kNeedU || kNeedB) {
        // the unary: at the side pitch for a chain-node belief, else at dst's
        constexpr bool kUs = EPI == kEpiUpNodeU;
        if (ty0 + kColTileY <= H) {
#pragma unroll
          for (int i = 0; i < kColRows; ++i) {
            const unsigned eb = kUs ? ebs0 + (unsigned)i * Wb2 : eb0 + (unsigned)i * Wb;
            if (kNeedU) pu[i] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(dun) + eb);
          }
        } else {
#pragma unroll
          for (int i = 0; i < kColRows; ++i) {
            const unsigned eb = (unsigned)(min(y0 + i, H - 1) * (kUs ? pb : pa) + xo) * 4u;
            if (kNeedU) pu[i] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(dun) + eb);
          }
        }
      }
      // the next tile's new rows (in the last tile: clamped rows, unused)
      const int s1 = ty0 + kColTileY + oy - ky + Lp;
      float pre[kNew];
#pragma unroll
      for (int k = 0; k < kNew; ++k) pre[k] = src_raw(s1 + wave + kColWaves * k);

      const float *tw = tile + wave * kColRows * kColTileX + lane;
      float a[kColRows], buf[kColRows];
#pragma unroll
      for (int i = 0; i < kColRows; ++i) {
        a[i] = 0.0f;
        buf[i] = tw[i * kColTileX];
      }
      // taps in runs of the ring's period, then a tail (the tap count is
      // padded to a multiple of 4 with zero taps)
      int t0 = 0;
      if constexpr (ROWS == 16) {
        for (; t0 + 16 <= Lp; t0 += 16) {
          col_chunk<0, 8, ROWS>(a, buf, fy, tw, t0);
          col_chunk<8, 8, ROWS>(a, buf, fy, tw, t0 + 8);
        }
        const int rem = Lp - t0;
        if (rem >= 8) {
          col_chunk<0, 8, ROWS>(a, buf, fy, tw, t0);
          if (rem == 12) col_chunk<8, 4, ROWS>(a, buf, fy, tw, t0 + 8);
        } else if (rem == 4) {
          col_chunk<0, 4, ROWS>(a, buf, fy, tw, t0);
        }
      } else {
        for (; t0 + 8 <= Lp; t0 += 8) col_chunk<0, 8, ROWS>(a, buf, fy, tw, t0);
        if (Lp - t0 == 4) col_chunk<0, 4, ROWS>(a, buf, fy, tw, t0);
      }

      if (kNeedB) {
#pragma unroll
        for (int i = 0; i < kColRows; ++i) {
          const unsigned eb = (unsigned)(min(y0 + i, H - 1) * pb + xo) * 4u;
          pbel[i] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(bel) + eb);
        }
      }
#pragma unroll
      for (int i = 0; i < kColRows; ++i) {
        const int y = y0 + i;
        // log, +M (:432-435); inverse warp (:437-448): outside -> LOG_ZERO
        const float msg = (col_ok && (unsigned)(y + oy) < (unsigned)H) ? log0(a[i]) + M : kLogZero;
        if (xin && y < yhi)
          acc.store(dst, dst2, eb0 + (unsigned)i * Wb, ebs0 + (unsigned)i * Wb2, msg, kNeedU ? pu[i] : 0.0f,
                    kNeedB ? pbel[i] : 0.0f, xmask, leaf_u);
      }
      if (more) {
        // move rows [64, nrows) to [0, Lp), then store the prefetched rows
        for (int c0 = 0; c0 < Lp; c0 += kColMove * kColWaves) {
          float mv[kColMove];
          __syncthreads();  // every wave is done with this tile
#pragma unroll
          for (int k = 0; k < kColMove; ++k) {
            const int rr = c0 + wave + kColWaves * k;
            if (rr < Lp) mv[k] = tile[(kColTileY + rr) * kColTileX + lane];
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < kColMove; ++k) {
            const int rr = c0 + wave + kColWaves * k;
            if (rr < Lp) tile[rr * kColTileX + lane] = mv[k];
          }
        }
#pragma unroll
        for (int k = 0; k < kNew; ++k) {
          const int row = s1 + wave + kColWaves * k;
          tile[(Lp + wave + kColWaves * k) * kColTileX + lane] = src_ok(row) ? pre[k] : 0.0f;
        }
        __syncthreads();
      }
    }
  }
  acc.flush(b, J, (long long)r * H * pa, lane);
}

// HI: the highest destination kind among the launch's jobs (the kernel holds
// only the epilogues up to it: 102-127 VGPRs below the five-grid downward
// one).  At most 128 VGPRs (4 waves per SIMD instead of 3): the five-grid
// epilogue fits without spilling vector registers once its stored belief is
// loaded after the filter (2.16 -> 1.33 ms per launch on 32 300x300x24
// grids; unbounded, at 156 VGPRs and 3 waves per SIMD, it takes 1.65 ms).
template <int HI, int ROWS, bool SPLIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_colpass(
    Bases b, const MsgJob *__restrict__ jobs, const float *__restrict__ taps, const int *__restrict__ shifts, int R,
    int xtiles) {
  extern __shared__ float4 lds4[];
  float *tile = reinterpret_cast<float *>(lds4);
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int r = (int)bx / xtiles;
  const int xt = (int)bx - r * xtiles;
  if (xt * kColTileX >= J.W) return;
  const int kind = epi_kind(J);
  if (HI == kEpiUpRoot || kind == kEpiUpRoot)
    colpass_strip<kEpiUpRoot, ROWS, SPLIT>(b, J, taps, shifts, R, r, xt, tile);
  else if (HI == kEpiUpNode || kind == kEpiUpNode)
    colpass_strip<kEpiUpNode, ROWS, SPLIT>(b, J, taps, shifts, R, r, xt, tile);
  else if (HI == kEpiUpNodeU || kind == kEpiUpNodeU)
    colpass_strip<kEpiUpNodeU, ROWS, SPLIT>(b, J, taps, shifts, R, r, xt, tile);
  else if (HI == kEpiDownLeaf || kind == kEpiDownLeaf)
    colpass_strip<kEpiDownLeaf, ROWS, SPLIT>(b, J, taps, shifts, R, r, xt, tile);
  else
    colpass_strip<kEpiDownMid, ROWS, SPLIT>(b, J, taps, shifts, R, r, xt, tile);
}

// ----------------------------------------------------------------- k_root

// KC: the number of root children when it is at most 8 (fully unrolled),
// else kMaxRootChildren with the count read at run time.
// LSE64: the root posterior's log-sum-exp over rotations accumulated in FP64
// with a running maximum shift (the BF16 methods); otherwise in FP32 with
// no shift, as the reference sums expf (:713-726: a belief above ~88.7
// overflows to +inf there and here)
template <int KC, bool LSE64>
__device__ __forceinline__ void root_pixels(const Bases &b, const RootJob &J, int R, int pix0) {
  const int H = J.H, W = J.W;
  const int HW = H * W;
  const int pix = pix0 + threadIdx.x;
  const bool ok = pix < HW;
  const int pc = ok ? pix : HW - 1;  // loads of lanes past the end read the last pixel
  const int yc = pc / W;
  const int x = pc - yc * W;
  // the lanes walk the dense grids (unary, marginal, log_from_root, root
  // posterior: index r HW + pix), so each wave writes its rows' lines whole;
  // the upward messages are read from the plan's slots (row pitch pm)
  const int pm = J.msg[0].pitch;
  const int HPm = H * pm;
  const int pcm = yc * pm + x;
  const int K = KC <= 8 ? KC : J.K;
  const float *un = cptr(b, J.unary);
  float *marg = mptr(b, J.marginal);

  unsigned kmax[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) kmax[k] = 0u;
  unsigned long long amax = 0ull;
  unsigned second = 0u;  // second largest root belief (key), for the MAP certificate
  float sum = 0.0f;
  bool first = true;
  double sum_d = 0.0;           // LSE64: sum of exp(belief - mref)
  float mref = -INFINITY;       //        the largest belief so far

  // the loads of slice r + 1 are issued before slice r is processed (its
  // stores could alias them as far as the compiler knows)
  float m[KC], u;
  auto load = [&](int r, float (&mm)[KC], float &uu) {
    const int e = r * HW + pc, em = r * HPm + pcm;
#pragma unroll
    for (int k = 0; k < KC; ++k) mm[k] = k < K ? cptr(b, J.msg[k])[em] : 0.0f;
    uu = un[e];
  };
  load(0, m, u);
  for (int r = 0; r < R; ++r) {
    float mn[KC], un_next;
    load(r + 1 < R ? r + 1 : r, mn, un_next);
    const int e = r * HW + pix;
    const float uu = masked(J.mask, r, x, W) ? kLogZero : u;
    // root belief: messages in joint order (:637), then the unary (:652-654)
    float bel = 0.0f;
#pragma unroll
    for (int k = 0; k < KC; ++k)
      if (k < K) bel = bel + m[k];
    if (J.add_unary) bel = bel + uu;
    // the root's rows are written whole by consecutive lanes: its stores
    // stream past the caches (nontemporal), 10 % faster than cached stores
    if (ok) {
      __builtin_nontemporal_store(bel, marg + e);
      const unsigned long long ak = amax_key(bel, (unsigned)e);
      const unsigned kv = (unsigned)(ak >> 32), kb = (unsigned)(amax >> 32);
      second = kv > kb ? kb : max(second, kv);
      amax = ak > amax ? ak : amax;
    }
    // log_from_root of child i: the other messages in order (:641-649), then
    // the root unary (:169)
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      if (i < K) {
        float fr = 0.0f;
#pragma unroll
        for (int k = 0; k < KC; ++k)
          if (k < K && k != i) fr = fr + m[k];
        fr = fr + uu;
        if (ok) {
          __builtin_nontemporal_store(fr, mptr(b, J.fr[i]) + e);
          kmax[i] = max(kmax[i], f2key(fr));
        }
      }
    }
    // root posterior over the valid rotations (:713-726)
    if ((J.valid_rot >> r) & 1ull) {
      if (LSE64) {
        if (bel > mref) {
          sum_d *= (double)__builtin_amdgcn_exp2f((mref - bel) * kLog2eF);
          mref = bel;
        }
        sum_d += (double)__builtin_amdgcn_exp2f((bel - mref) * kLog2eF);
      } else {
        const float ev = expf(bel);
        sum = first ? ev : sum + ev;
        first = false;
      }
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) m[k] = mn[k];
    u = un_next;
  }
  if (ok && J.root_post.buf >= 0 && b.root)
    mptr(b, J.root_post)[pix] = LSE64 ? (sum_d > 0.0 ? log0((float)sum_d) + mref : kLogZero) : log0(sum);

  const int lane = threadIdx.x & 63;
  if (b.top2_keys) top2_flush(b.top2_keys + J.map_slot, (unsigned)(amax >> 32), second, lane);
  amax = wave_max_u64(amax);
  if (lane == 0 && amax) atomicMax(b.map_keys + J.map_slot, amax);
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    if (k < K) {
      unsigned v = wave_max_u32(kmax[k]);
      if (lane == 0 && v) atomicMax(b.max_keys + J.fr_key[k], v);
    }
  }
}

// one instantiation per child count, chosen on the host: the variant for a
// small tree then holds only its own K message registers (46 VGPRs for K = 5,
// 8 waves per SIMD), not those of the widest tree
template <int KC, bool LSE64>
__global__ __launch_bounds__(256) void k_root(Bases b, const RootJob *__restrict__ jobs, int R) {
  unsigned bx, by;
  xcd_block(bx, by);
  const RootJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int pix0 = (int)bx * blockDim.x;
  if (pix0 >= J.H * J.W) return;
  root_pixels<KC, LSE64>(b, J, R, pix0);
}

// ------------------------------------------------------------ k_map_decode

__global__ void k_map_decode(Bases b, const MapJob *__restrict__ jobs, int n, ps_map *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapJob J = jobs[i];
  if (skip_inst(b, J.inst)) return;
  const unsigned long long k = b.map_keys[J.slot];
  ps_map m;
  if (k == 0ull) {
    m.rotidx = -1;
    m.y = -1;
    m.x = -1;
    m.score = -INFINITY;
  } else {
    const unsigned idx = 0xffffffffu - (unsigned)(k & 0xffffffffull);
    const unsigned hw = (unsigned)J.H * (unsigned)J.W;
    m.rotidx = (int)(idx / hw);
    const unsigned rem = idx - (unsigned)m.rotidx * hw;
    m.y = (int)(rem / (unsigned)J.W);
    m.x = (int)(rem - (unsigned)m.y * (unsigned)J.W);
    m.score = key2f((unsigned)(k >> 32));
  }
  out[J.out] = m;
}

// ------------------------------------------------------------- k_certify
// MAP certificate of one (instance, scale, part): with every log-marginal
// value off by at most eps, the computed argmax is the exact one whenever
// the computed best beats the second best by more than tau = 2 eps.  An
// instance with any slot at or below tau is flagged for re-evaluation.

__global__ void k_certify(const unsigned long long *__restrict__ top2, int n_map, int map_per_inst, float tau,
                          unsigned *flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_map) return;
  const unsigned long long t = top2[i];
  const unsigned k1 = (unsigned)(t >> 32), k2 = (unsigned)t;
  const float v1 = k1 ? key2f(k1) : -INFINITY;
  const float v2 = k2 ? key2f(k2) : -INFINITY;
  if (!(v1 - v2 > tau)) flags[i / map_per_inst] = 1u;
}

__global__ void k_invert_flags(const unsigned *__restrict__ in, unsigned *__restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] ? 0u : 1u;
}

// ------------------------------------------------------------ host launchers

hipError_t launch_invert_flags(const unsigned *in, unsigned *out, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_invert_flags, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_certify(const unsigned long long *top2, int n_map, int n_inst, float tau, unsigned *flags,
                          hipStream_t s) {
  if (n_map <= 0) return hipSuccess;
  const int mpi = std::max(1, n_map / std::max(1, n_inst));
  hipLaunchKernelGGL(k_certify, dim3((n_map + 255) / 256), dim3(256), 0, s, top2, n_map, mpi, tau, flags);
  return hipGetLastError();
}

hipError_t launch_reset(const Bases &b, int n_keys, int n_map, int n_inst, unsigned *flags, int n_flags,
                        hipStream_t s) {
  int n = std::max(std::max(n_keys, n_map), flags ? n_flags : 0);
  if (n <= 0) return hipSuccess;
  const int kpi = std::max(1, n_keys / std::max(1, n_inst)), mpi = std::max(1, n_map / std::max(1, n_inst));
  hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, s, b, n_keys, n_map, kpi, mpi, flags, n_flags);
  return hipGetLastError();
}

hipError_t launch_max(const Bases &b, const MaxJob *jobs, int n_jobs, int R, int max_hw, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  const int chunk = 16384;  // elements per block, a multiple of 4
  const long long n = (long long)R * max_hw;
  dim3 grid((unsigned)((n + chunk - 1) / chunk), n_jobs);
  hipLaunchKernelGGL(k_max, grid, dim3(256), 0, s, b, jobs, R, chunk);
  return hipGetLastError();
}

static bool rowpass_fixed(int R) { return R == 24 || R == 48; }

size_t rowpass_lds_bytes(int R, int W, int kx) {
  RowGeom g = row_geom(W, kx);
  // the fixed-R kernels stage one slice group at a time; the generic one all
  // slices and its E buffer
  size_t p = rowpass_fixed(R) ? (size_t)(R / row_groups(R)) * g.Pw : (size_t)R * g.Pw + (size_t)R * ((W + 3) & ~3);
  return sizeof(float) * p;
}

// The dynamic-LDS attribute of a kernel is one process-wide value per device
// and function, while plans with different needs coexist: it is only ever
// raised (to the largest size asked for so far), under a lock, so a later
// plan never lowers what an earlier plan's launches rely on.
hipError_t raise_lds_limit(const void *f, size_t bytes) {
  if (bytes <= 64 * 1024) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> high;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  size_t &h = high[std::make_pair(dev, f)];
  if (bytes <= h) return hipSuccess;
  e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) h = bytes;
  return e;
}

hipError_t rowpass_set_lds_limit(size_t bytes) {
  hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_r<24, false>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_r<48, false>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_r<24, true>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_r<48, true>), bytes);
  if (e != hipSuccess) return e;
  return raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_any), bytes);
}

template <int ROWS, bool SPLIT>
static hipError_t colpass_rows_set_lds_limit(size_t bytes) {
  hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass<kEpiUpRoot, ROWS, SPLIT>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass<kEpiUpNode, ROWS, SPLIT>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass<kEpiUpNodeU, ROWS, SPLIT>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass<kEpiDownLeaf, ROWS, SPLIT>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass<kEpiDownMid, ROWS, SPLIT>), bytes);
  return e;
}

// bytes: the LDS of the 16-row tile (the larger one)
hipError_t colpass_set_lds_limit(size_t bytes) {
  hipError_t e = colpass_rows_set_lds_limit<8, false>(bytes);
  if (e == hipSuccess) e = colpass_rows_set_lds_limit<16, false>(bytes);
  if (e == hipSuccess) e = colpass_rows_set_lds_limit<8, true>(bytes);
  return e;
}

// threads: block size (a multiple of 64; 0 = one thread per column, at most
// 512); every loop of the row pass strides by the block size
// split: some job's warp changes its offset along an axis (MsgJob::split)
hipError_t launch_rowpass(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                          int R, int max_h, int max_w, size_t lds_bytes, hipStream_t s, int threads, bool split) {
  if (n_jobs <= 0) return hipSuccess;
  dim3 grid(max_h, n_jobs);
  const int nt = threads > 0 ? std::min(threads, std::max(64, (max_w + 63) / 64 * 64))
                             : std::min(512, (max_w + 63) / 64 * 64);
  if (R == 24 && !split)
    hipLaunchKernelGGL((k_rowpass_r<24, false>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts);
  else if (R == 24)
    hipLaunchKernelGGL((k_rowpass_r<24, true>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts);
  else if (R == 48 && !split)  // the reference's default grid
    hipLaunchKernelGGL((k_rowpass_r<48, false>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts);
  else if (R == 48)
    hipLaunchKernelGGL((k_rowpass_r<48, true>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts);
  else
    hipLaunchKernelGGL(k_rowpass_any, grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, R);
  return hipGetLastError();
}

// LDS of a column-pass tile: rows 0 = the larger (16-row) tile
size_t colpass_lds_bytes(int ky, int rows) {
  return sizeof(float) * (size_t)col_tile_rows(ky, rows > 0 ? rows : 16) * kColTileX;
}

template <int ROWS, bool SPLIT>
static void launch_colpass_rows(int epi_hi, dim3 grid, size_t lds_bytes, hipStream_t s, const Bases &b,
                                const MsgJob *jobs, const float *taps, const int *shifts, int R, int xtiles) {
  const dim3 blk(kColWaves * 64);
  switch (epi_hi) {
    case kEpiUpRoot: hipLaunchKernelGGL((k_colpass<kEpiUpRoot, ROWS, SPLIT>), grid, blk, lds_bytes, s, b, jobs, taps, shifts, R, xtiles); break;
    case kEpiUpNode: hipLaunchKernelGGL((k_colpass<kEpiUpNode, ROWS, SPLIT>), grid, blk, lds_bytes, s, b, jobs, taps, shifts, R, xtiles); break;
    case kEpiUpNodeU: hipLaunchKernelGGL((k_colpass<kEpiUpNodeU, ROWS, SPLIT>), grid, blk, lds_bytes, s, b, jobs, taps, shifts, R, xtiles); break;
    case kEpiDownLeaf: hipLaunchKernelGGL((k_colpass<kEpiDownLeaf, ROWS, SPLIT>), grid, blk, lds_bytes, s, b, jobs, taps, shifts, R, xtiles); break;
    default: hipLaunchKernelGGL((k_colpass<kEpiDownMid, ROWS, SPLIT>), grid, blk, lds_bytes, s, b, jobs, taps, shifts, R, xtiles); break;
  }
}

// epi_hi: the highest destination kind (EpiKind) among the jobs; rows:
// output rows per lane (8 or 16; 0 = the default); lds_bytes_16: the LDS of
// the 16-row tile for the jobs' widest y filter (the 8-row tile takes 32
// rows of 256 bytes less)
// (split: some job's warp changes its offset along an axis; those launches
// take 8 rows per lane)
hipError_t launch_colpass(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                          int R, int epi_hi, int max_w, size_t lds_bytes_16, hipStream_t s, int rows, bool split) {
  if (n_jobs <= 0) return hipSuccess;
  if (rows == 0 || split) rows = kColRowsDefault;
  if (epi_hi < kEpiUpRoot || epi_hi > kEpiDownMid || (rows != 8 && rows != 16)) return hipErrorInvalidValue;
  const int xtiles = (max_w + kColTileX - 1) / kColTileX;
  dim3 grid(R * xtiles, n_jobs);
  const size_t lds_8 = lds_bytes_16 - sizeof(float) * kColWaves * 8 * kColTileX;
  if (split)
    launch_colpass_rows<8, true>(epi_hi, grid, lds_8, s, b, jobs, taps, shifts, R, xtiles);
  else if (rows == 16)
    launch_colpass_rows<16, false>(epi_hi, grid, lds_bytes_16, s, b, jobs, taps, shifts, R, xtiles);
  else
    launch_colpass_rows<8, false>(epi_hi, grid, lds_8, s, b, jobs, taps, shifts, R, xtiles);
  return hipGetLastError();
}

// k: the root's child count, the same for every job (one tree per plan); a
// root without children and counts above 8 take the variant that reads J.K
template <bool LSE64>
static void launch_root_k(const Bases &b, const RootJob *jobs, int k, int R, dim3 grid, dim3 blk, hipStream_t s) {
  switch (k) {
    case 1: hipLaunchKernelGGL((k_root<1, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 2: hipLaunchKernelGGL((k_root<2, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 3: hipLaunchKernelGGL((k_root<3, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 4: hipLaunchKernelGGL((k_root<4, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 5: hipLaunchKernelGGL((k_root<5, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 6: hipLaunchKernelGGL((k_root<6, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 7: hipLaunchKernelGGL((k_root<7, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    case 8: hipLaunchKernelGGL((k_root<8, LSE64>), grid, blk, 0, s, b, jobs, R); break;
    default: hipLaunchKernelGGL((k_root<kMaxRootChildren, LSE64>), grid, blk, 0, s, b, jobs, R); break;
  }
}

hipError_t launch_root(const Bases &b, const RootJob *jobs, int n_jobs, int k, int R, int max_hw, bool lse64,
                       hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  if (k < 0 || k > kMaxRootChildren) return hipErrorInvalidValue;
  dim3 grid((max_hw + 255) / 256, n_jobs);
  const dim3 blk(256);
  if (lse64)
    launch_root_k<true>(b, jobs, k, R, grid, blk, s);
  else
    launch_root_k<false>(b, jobs, k, R, grid, blk, s);
  return hipGetLastError();
}

hipError_t launch_map_decode(const Bases &b, const MapJob *jobs, int n, ps_map *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_map_decode, dim3((n + 255) / 256), dim3(256), 0, s, b, jobs, n, out);
  return hipGetLastError();
}

}  // namespace ps
