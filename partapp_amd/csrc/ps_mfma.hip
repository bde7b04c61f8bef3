// BF16 banded-Toeplitz variant of a message's spatial filter on the gfx950
// matrix cores.  Same data flow as the FP32 row / column passes (the forward
// shift, exp, rotation filter, log, inverse shift and the caller's
// accumulation are unchanged); only the two truncated Gaussian passes
// (multi_array_filter.hpp:277-288 along x, :304-317 along y) become GEMMs
// against a Toeplitz matrix of the taps:
//
//   out[m][n] = sum_k A[m][k] B[k][n],  band entries f[k - m] (y pass, A) or
//                                       f[k - n] (x pass, B), zero elsewhere
//
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  A 32-wide output tile
// needs K = 32 + 2k inputs, rounded up to KK steps of 16.  TERMS selects the
// precision: 1 multiplies bf16(data) by bf16(taps); 2 splits the data into
// bf16 hi + lo parts (data to about 16 significant bits, taps in bf16); 3
// splits both and adds hi*hi + lo*hi + hi*lo.  The column pass rescales every
// column of a slice by a power of two taken from its first rows (the y filter
// never mixes columns), so the bf16 operands keep their range; the scale is
// undone exactly after the GEMM.  The exponential of the row pass and the
// logarithm of the column pass use the hardware instructions (a few ulp).
//
// Operand layouts of the 32x32x16 bf16 MFMA (lane l, t = 0..7):
//   A: A[l % 32][(l / 32) * 8 + t]   B: B[(l / 32) * 8 + t][l % 32]
//   C: acc[i] = C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32]
// Both Toeplitz operands therefore hold f[16 kk + 8 (l / 32) + t - l % 32]
// in lane l at step kk: one fragment table per filter, built on the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ps_device.hpp"
#include "ps_internal.hpp"
#include "ps_rowgather.hpp"

namespace ps {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ unsigned short bf16_bits(float v) {
  return __builtin_bit_cast(unsigned short, (__bf16)v);
}
__device__ __forceinline__ bf16x8_t ld_frag(const unsigned short *p) {
  return *reinterpret_cast<const bf16x8_t *>(p);
}
__device__ __forceinline__ bf16x8_t ld_frag(const uint4 *p) { return *reinterpret_cast<const bf16x8_t *>(p); }
__device__ __forceinline__ f32x16_t mfma(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// eight floats as bf16 hi parts and lo parts (v - float(hi) is exact)
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8_t &hi, bf16x8_t &lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const __bf16 h = (__bf16)v[t];
    hi[t] = h;
    lo[t] = (__bf16)(v[t] - (float)h);
  }
}

// 2^-e and 2^e that bring the maximum mx into [0.5, 1), with e limited to
// [-60, 60] (both 1 when mx is 0 or not finite)
__device__ __forceinline__ void slice_scale(float mx, float &scale, float &unscale) {
  int e = 0;
  if (mx > 0.0f && mx < INFINITY) (void)frexpf(mx, &e);
  e = max(-60, min(60, e));
  scale = ldexpf(1.0f, -e);
  unscale = ldexpf(1.0f, e);
}

}  // namespace

// bf16 per LDS row of the row pass: the last 32-wide tile's window end (and
// the kx + W written columns), with a row stride of 16 bytes modulo 32 so the
// 16-byte operand reads of 32 different rows hit distinct banks
__host__ __device__ __forceinline__ int mfma_row_pitch(int W, int kx, int KK) {
  int need = 32 * ((W + 31) / 32 - 1) + 16 * KK;
  need = need > kx + W ? need : kx + W;
  return (need + 7) / 16 * 16 + 8;
}
// ring rows of the column pass and their pitch (bf16 per column, 16 bytes
// modulo 32 bytes: 16-byte accesses of 64 columns are conflict free)
__host__ __device__ __forceinline__ int mfma_ring_rows(int KK) { return 32 + 16 * KK; }
__host__ __device__ __forceinline__ int mfma_ring_pitch(int KK) { return mfma_ring_rows(KK) + 8; }

// ------------------------------------------------------- row pass (x filter)
// One block per (message, run of `rows` output rows), one lane per column
// (W <= blockDim.x).  For each row y: the lane's column of all R slices
// (shift, exp, rotation filter) is stored split into bf16 hi / lo rows
// P[r][kx + x] in LDS, while the loads of the next row are already in
// flight; then each wave takes 32-wide column tiles and runs the KK-step GEMM
// P (R x K, rows clamped to R - 1 beyond R) times the Toeplitz fragments, one
// accumulator for hi*hi and one for the two cross terms.  The output of slice
// r is stored shifted by the inverse warp's column offset (column x at
// x - inv_ox[r]), so the column pass reads 64-column strips aligned with its
// outputs; columns that no output reads are left unwritten.  The row pass
// works on exp(b - M) as the FP32 pass does (M the source grid's maximum).
template <int R, int TERMS>
__global__ __launch_bounds__(512) void k_rowpass_mfma(Bases b, const MsgJob *__restrict__ jobs,
                                                      const float *__restrict__ taps, const int *__restrict__ shifts,
                                                      const uint4 *__restrict__ frags, int rows) {
  static_assert(R <= 32, "one 32-row MFMA tile holds every rotation slice");
  extern __shared__ uint4 lds[];
  const MsgJob &J = jobs[blockIdx.y];
  const int H = J.H, W = J.W;
  const int y0 = blockIdx.x * rows;
  if (y0 >= H) return;
  const int y1 = min(H, y0 + rows);
  const int kx = J.kx, KK = J.kk_x;
  const int nf = TERMS == 3 ? 2 : 1;  // tap parts
  const int nd = TERMS >= 2 ? 2 : 1;  // data parts
  const int pitch = mfma_row_pitch(W, kx, KK);
  uint4 *fr = lds;  // [nf][KK][64] Toeplitz fragments
  unsigned short *Ph = reinterpret_cast<unsigned short *>(fr + nf * KK * 64);
  unsigned short *Pl = Ph + R * pitch;
  int *oxs = reinterpret_cast<int *>(Ph + nd * R * pitch);  // [R] inverse column shifts

  const uint4 *gf = frags + J.frag_x;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];
  if (threadIdx.x < R) oxs[threadIdx.x] = shifts[J.shifts + 2 * R + threadIdx.x];
  // the halo columns stay zero for every row of the block
  const int per_row = pitch - W;
  for (int i = threadIdx.x; i < R * per_row; i += blockDim.x) {
    const int r = i / per_row;
    const int c = i - r * per_row;
    const int col = r * pitch + (c < kx ? c : c + W);
    Ph[col] = 0;
    if (TERMS >= 2) Pl[col] = 0;
  }

  const float M = key2f(b.max_keys[J.src_key]);
  const float negM = -M;
  const float e_fill = expf(kLogZero + negM);
  const float *src = cptr(b, J.src);
  const int *oxo = shifts + J.shifts + 4 * R;
  const int *oyo = oxo + R;
  const float *fr_rot = taps + J.taps_r;
  const bool zero_src = J.src_kind == kSrcZero;
  const UnaryMask mk = J.src_mask;
  const bool use_mask = J.src_kind == kSrcUnary && (mk.rot_bits != 0ull || mk.strip != 0);
  const int HW = H * W;  // R * H * W < 2^31 is checked when the plan is built
  const int tp = J.tpitch;
  const int tplane = H * tp;
  const int x = threadIdx.x;
  const bool has_x = x < W;
  const int xl = has_x ? x : 0;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int m = min(lane & 31, R - 1);
  const int kq = (lane >> 5) << 3;
  const int ntile = (W + 31) >> 5;

  // the loads of row y + 1 are issued once row y is in LDS, and are in flight
  // during row y's GEMM
  float v[R];
  unsigned long long ok = row_loads<R>(src, oxo, oyo, H, W, HW, J.delta, zero_src, y0, xl, v);
  for (int y = y0; y < y1; ++y) {
    if (has_x) {
      float o[R];
      row_finish<R, true>(v, ok, oxo, fr_rot, negM, e_fill, W, J.delta, J.rot_mode, J.rot_len, use_mask, mk, x, o);
#pragma unroll
      for (int ro = 0; ro < R; ++ro) Ph[ro * pitch + kx + x] = bf16_bits(o[ro]);
      if (TERMS >= 2) {
#pragma unroll
        for (int ro = 0; ro < R; ++ro) Pl[ro * pitch + kx + x] = bf16_bits(o[ro] - (float)(__bf16)o[ro]);
      }
    }
    if (y + 1 < y1) ok = row_loads<R>(src, oxo, oyo, H, W, HW, J.delta, zero_src, y + 1, xl, v);
    __syncthreads();
    float *T = mptr(b, J.tmp) + y * tp;
    for (int t = wave; t < ntile; t += nw) {
      const unsigned short *ah = Ph + m * pitch + (t << 5) + kq;
      const unsigned short *al = Pl + m * pitch + (t << 5) + kq;
      f32x16_t c0 = {}, c1 = {};
      for (int kk = 0; kk < KK; ++kk) {
        const bf16x8_t a_hi = ld_frag(ah + 16 * kk);
        const bf16x8_t b_hi = ld_frag(fr + kk * 64 + lane);
        c0 = mfma(a_hi, b_hi, c0);
        if (TERMS >= 2) c1 = mfma(ld_frag(al + 16 * kk), b_hi, c1);
        if (TERMS == 3) c1 = mfma(a_hi, ld_frag(fr + (KK + kk) * 64 + lane), c1);
      }
      const int xo = (t << 5) + (lane & 31);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int r = 8 * (i >> 2) + ((lane >> 5) << 2) + (i & 3);
        const float v = TERMS >= 2 ? c0[i] + c1[i] : c0[i];
        const int xs = xo - oxs[r < R ? r : 0];
        if (r < R && xo < W && xs >= 0 && xs < W) T[r * tplane + xs] = v;
      }
    }
    __syncthreads();  // P is rewritten by the next row
  }
}

// ---------------------------------------------------- column pass (y filter)
// One block per (message, slice r, strip of 64 columns), walking down the
// strip in tiles of 64 output rows.  The row pass stored slice r shifted by
// the inverse column offset, so the strip's columns are the outputs' own.
// The source rows a tile needs (32 + 16 KK: the wave's 32 rows plus the
// band), scaled by the slice's power of two and split into bf16 hi / lo, sit
// in an LDS ring with one column per run of rows (the GEMM operand reads 8
// rows of one column per lane; staging writes 16 rows of one column per
// thread); the next tile reuses all but 64 of them.  Each wave computes a
// 32 x 32 quarter of the tile, then the epilogue of the FP32 column pass with
// its operands loaded before the GEMM.
template <int TERMS>
__global__ __launch_bounds__(256) void k_colpass_mfma(Bases b, const MsgJob *__restrict__ jobs,
                                                      const uint4 *__restrict__ frags, const int *__restrict__ shifts,
                                                      int R, int xtiles) {
  extern __shared__ uint4 lds[];
  const MsgJob &J = jobs[blockIdx.y];
  const int H = J.H, W = J.W;
  const int r = blockIdx.x / xtiles;
  const int xt = blockIdx.x - r * xtiles;
  if (xt * 64 >= W) return;
  const int KK = J.kk_y;
  const int nrows = mfma_ring_rows(KK);
  const int cs = mfma_ring_pitch(KK);
  const int nf = TERMS == 3 ? 2 : 1;  // tap parts
  uint4 *fr = lds;  // [nf][KK][64] Toeplitz fragments
  unsigned short *Rh = reinterpret_cast<unsigned short *>(lds + nf * KK * 64);
  unsigned short *Rl = Rh + 64 * cs;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint4 *gf = frags + J.frag_y;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];
  const int *sh = shifts + J.shifts;
  const int ox = sh[2 * R + r];
  const int oy = sh[3 * R + r];
  const int ky = J.ky;
  const int HW = H * W;
  float *cscale = reinterpret_cast<float *>(Rl + (TERMS >= 2 ? 64 * cs : 0));  // [64] 2^e of each column
  unsigned *cmax = reinterpret_cast<unsigned *>(cscale + 64);                   // [64] max of the first rows
  if (threadIdx.x < 64) cmax[threadIdx.x] = 0u;
  const float M = key2f(b.max_keys[J.src_key]);
  float *dst = mptr(b, J.dst) + r * HW;
  const float *dun = (J.dst_unary.buf >= 0 ? cptr(b, J.dst_unary) : cptr(b, J.dst)) + r * HW;
  float *dst2 = J.dst2.buf >= 0 ? mptr(b, J.dst2) : nullptr;
  const bool need_u = (J.dst_kind == kDstUpNode && J.dst_add_unary) || J.dst_kind == kDstDown;
  const bool need_b = J.dst_kind == kDstDown && J.dst_base == 0;

  // staging: this thread loads column xt * 64 + lane of the shifted slice, 16
  // consecutive rows at a time (rows outside the grid are zero)
  const int lcol = xt * 64 + lane;
  const int tp = J.tpitch;
  const float *Tc = cptr(b, J.tmp) + (long long)r * H * tp + (lcol < W ? lcol : 0);
  const bool lcol_ok = lcol < W;
  float scale = 1.0f;
  auto load16 = [&](int row0, float (&v)[16]) {
    const bool all_in = lcol_ok && row0 >= 0 && row0 + 16 <= H;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = row0 + k;
      const bool in = all_in || (lcol_ok && row >= 0 && row < H);
      const float t = Tc[(in ? row : 0) * tp];
      v[k] = in ? t : 0.0f;
    }
  };
  auto put16 = [&](int prow, const float (&v)[16]) {
    float w[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = v[k] * scale;
    bf16x8_t h0, l0, h1, l1;
    split8(*reinterpret_cast<const float(*)[8]>(&w[0]), h0, l0);
    split8(*reinterpret_cast<const float(*)[8]>(&w[8]), h1, l1);
    *reinterpret_cast<bf16x8_t *>(Rh + lane * cs + prow) = h0;
    *reinterpret_cast<bf16x8_t *>(Rh + lane * cs + prow + 8) = h1;
    if (TERMS >= 2) {
      *reinterpret_cast<bf16x8_t *>(Rl + lane * cs + prow) = l0;
      *reinterpret_cast<bf16x8_t *>(Rl + lane * cs + prow + 8) = l1;
    }
  };
  // tile rows are source rows ty0 + oy - ky + [0, nrows), at ring row (base +
  // row) mod nrows.  The y filter never mixes columns, so every column gets its
  // own power-of-two scale, from the maximum of its first nrows rows (capped so
  // that later, larger rows cannot overflow).
  constexpr int kFill = 3;  // chunks of 16 rows per thread in the first fill (nrows <= 160)
  float fill[kFill][16];
  float mx = 0.0f;
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) {
      load16(oy - ky + 16 * c, fill[q]);
#pragma unroll
      for (int k = 0; k < 16; ++k) mx = fmaxf(mx, fill[q][k]);
    }
  }
  __syncthreads();  // cmax is initialised
  if (mx > 0.0f) atomicMax(cmax + lane, __float_as_uint(mx));
  __syncthreads();
  {
    float sc, unsc;
    slice_scale(__uint_as_float(cmax[lane]), sc, unsc);
    scale = sc;
    if (wave == 0) cscale[lane] = unsc;
  }
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) put16(16 * c, fill[q]);
  }
  __syncthreads();

  const int cx = (wave & 1) * 32, ry = (wave >> 1) * 32;
  const int n = lane & 31;
  const int kq = (lane >> 5) << 3;
  const int x = xt * 64 + cx + n;
  const bool xin = x < W;
  const bool col_ok = xin && x + ox >= 0 && x + ox < W;
  const bool xmask = ((J.dst_mask.rot_bits >> r) & 1ull) || x < J.dst_mask.strip || x >= W - J.dst_mask.strip;
  const unsigned short *bh_col = Rh + (cx + n) * cs;
  const unsigned short *bl_col = Rl + (cx + n) * cs;
  const float unscale = cscale[cx + n];
  const int xo = xin ? x : 0;
  DstAcc da;
  int base = 0;
  for (int ty0 = 0; ty0 < H; ty0 += 64) {
    const bool more = ty0 + 64 < H;
    float pre[16];
    if (more) load16(ty0 + 64 + oy - ky + nrows - 64 + 16 * wave, pre);
    // epilogue operands of this lane's 16 outputs (rows y0 + 8 (i / 4) + i % 4)
    const int y0 = ty0 + ry + ((lane >> 5) << 2);
    float pu[16], pb[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int y = y0 + 8 * (i >> 2) + (i & 3);
      const int e = (y < H ? y : 0) * W + xo;
      pu[i] = need_u ? dun[e] : 0.0f;
      pb[i] = need_b ? dst[e] : 0.0f;
    }
    f32x16_t c0 = {}, c1 = {};
    int lrow = base + ry + kq;
    if (lrow >= nrows) lrow -= nrows;
    for (int kk = 0; kk < KK; ++kk) {
      const bf16x8_t a_hi = ld_frag(fr + kk * 64 + lane);
      const bf16x8_t b_hi = ld_frag(bh_col + lrow);
      c0 = mfma(a_hi, b_hi, c0);
      if (TERMS >= 2) c1 = mfma(a_hi, ld_frag(bl_col + lrow), c1);
      if (TERMS == 3) c1 = mfma(ld_frag(fr + (KK + kk) * 64 + lane), b_hi, c1);
      lrow += 16;
      if (lrow >= nrows) lrow -= nrows;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int y = y0 + 8 * (i >> 2) + (i & 3);
      if (!xin || y >= H) continue;
      const int ysrc = y + oy;
      const float v = TERMS >= 2 ? c0[i] + c1[i] : c0[i];
      const float msg = (col_ok && ysrc >= 0 && ysrc < H) ? log0_fast(v * unscale) + M : kLogZero;
      dst_store(J, dst, dst2 ? dst2 + r * HW : nullptr, y * W + x, r * HW + y * W + x, xmask, msg, pu[i], pb[i],
                da);
    }
    if (more) {
      __syncthreads();  // every wave is done with this tile's ring rows
      int prow = base + 16 * wave;
      if (prow >= nrows) prow -= nrows;
      put16(prow, pre);
      base += 64;
      if (base >= nrows) base -= nrows;
      __syncthreads();
    }
  }
  dst_flush(b, J, dst2 != nullptr, da, lane);
}

// ------------------------------------------------------------ host side

size_t rowpass_mfma_lds_bytes(int R, int W, int kx, int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1, nd = terms >= 2 ? 2 : 1;
  return 16 * (size_t)nf * KK * 64 + 2 * (size_t)nd * R * mfma_row_pitch(W, kx, KK) + 4 * (size_t)R;
}
size_t colpass_mfma_lds_bytes(int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1, nd = terms >= 2 ? 2 : 1;
  return 16 * (size_t)nf * KK * 64 + 2 * (size_t)nd * 64 * mfma_ring_pitch(KK) + 8 * 64;
}

hipError_t mfma_set_lds_limit(size_t row_bytes, size_t col_bytes) {
  auto set = [](const void *f, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  };
  hipError_t e = set(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 1>), row_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 2>), row_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 3>), row_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_colpass_mfma<1>), col_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_colpass_mfma<2>), col_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_colpass_mfma<3>), col_bytes);
  return e;
}

hipError_t launch_rowpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                               const uint4 *frags, int R, int terms, int rows, int max_h, int max_w,
                               size_t lds_bytes, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  // one lane per column: the plan only routes grids of width <= 512 here
  if (R != 24 || terms < 1 || terms > 3 || max_w > 512 || rows < 1) return hipErrorInvalidValue;
  dim3 grid((max_h + rows - 1) / rows, n_jobs);
  const int nt = std::min(512, std::max(64, (max_w + 63) / 64 * 64));
  if (terms == 3)
    hipLaunchKernelGGL((k_rowpass_mfma<24, 3>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, frags, rows);
  else if (terms == 2)
    hipLaunchKernelGGL((k_rowpass_mfma<24, 2>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, frags, rows);
  else
    hipLaunchKernelGGL((k_rowpass_mfma<24, 1>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, frags, rows);
  return hipGetLastError();
}

hipError_t launch_colpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const uint4 *frags, const int *shifts,
                               int R, int terms, int max_w, size_t lds_bytes, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  const int xtiles = (max_w + 63) / 64;
  dim3 grid(R * xtiles, n_jobs);
  if (terms == 3)
    hipLaunchKernelGGL(k_colpass_mfma<3>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  else if (terms == 2)
    hipLaunchKernelGGL(k_colpass_mfma<2>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  else
    hipLaunchKernelGGL(k_colpass_mfma<1>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  return hipGetLastError();
}

}  // namespace ps
