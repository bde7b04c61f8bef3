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
// precision: 1 multiplies bf16(data) by bf16(taps); 3 splits both into
// bf16 hi + lo parts and adds hi*hi + lo*hi + hi*lo (about 16 significant
// bits).  The column pass rescales every rotation slice by a power of two
// taken from its maximum (gathered by the row pass), so the bf16 operands
// keep their range; the scale is undone exactly after the GEMM.
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
// v = hi + lo to about 16 significant bits; v - float(hi) is exact
__device__ __forceinline__ void split_bf16(float v, unsigned short &hi, unsigned short &lo) {
  const __bf16 h = (__bf16)v;
  hi = __builtin_bit_cast(unsigned short, h);
  lo = bf16_bits(v - (float)h);
}
__device__ __forceinline__ bf16x8_t ld_frag(const unsigned short *p) {
  return *reinterpret_cast<const bf16x8_t *>(p);
}
__device__ __forceinline__ bf16x8_t ld_frag(const uint4 *p) { return *reinterpret_cast<const bf16x8_t *>(p); }
__device__ __forceinline__ f32x16_t mfma(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// 2^-e and 2^e for the slice maximum mx (both 1 when mx is 0 or not finite)
__device__ __forceinline__ void slice_scale(float mx, float &scale, float &unscale) {
  int e = 0;
  if (mx > 0.0f && mx < INFINITY) (void)frexpf(mx, &e);
  scale = ldexpf(1.0f, -e);
  unscale = ldexpf(1.0f, e);
}

}  // namespace

// bf16 elements per LDS row of the row pass: the last 32-wide tile's window
// end, padded so that the row stride is 16 bytes modulo 128 (the 16-byte
// operand reads of 32 rows then hit distinct banks)
__host__ __device__ __forceinline__ int mfma_row_pitch(int W, int KK) {
  const int need = 32 * ((W + 31) / 32 - 1) + 16 * KK;
  return (need + 55) / 64 * 64 + 8;
}
// ring rows of the column pass and their pitch (bf16 elements per column)
__host__ __device__ __forceinline__ int mfma_ring_rows(int KK) { return 32 + 16 * KK; }
__host__ __device__ __forceinline__ int mfma_ring_pitch(int KK) { return (mfma_ring_rows(KK) + 55) / 64 * 64 + 8; }

// ------------------------------------------------------- row pass (x filter)
// One block per (message, row y), as the FP32 row pass: each lane gathers its
// column of all R slices (shift, exp, rotation filter), splits it into bf16
// hi / lo rows P[r][kx + x] in LDS; then each wave takes 32-wide column tiles
// and runs the KK-step GEMM P (R x K, rows clamped to R - 1 beyond R) times
// the Toeplitz fragments.  Also reduces the maximum of every output slice
// into the job's slice keys.
template <int R, int TERMS>
__global__ __launch_bounds__(512) void k_rowpass_mfma(Bases b, const MsgJob *__restrict__ jobs,
                                                      const float *__restrict__ taps, const int *__restrict__ shifts,
                                                      const uint4 *__restrict__ frags) {
  static_assert(R <= 32, "one 32-row MFMA tile holds every rotation slice");
  extern __shared__ uint4 lds[];
  const MsgJob &J = jobs[blockIdx.y];
  const int y = blockIdx.x;
  const int H = J.H, W = J.W;
  if (y >= H) return;
  const int kx = J.kx, KK = J.kk_x;
  const int nf = TERMS == 3 ? 2 : 1;
  const int pitch = mfma_row_pitch(W, KK);
  uint4 *fr = lds;  // [nf][KK][64] fragments
  unsigned short *Ph = reinterpret_cast<unsigned short *>(lds + nf * KK * 64);
  unsigned short *Pl = Ph + R * pitch;
  unsigned *smax = reinterpret_cast<unsigned *>(Pl + (TERMS == 3 ? R * pitch : 0));  // [R] slice maxima

  const uint4 *gf = frags + J.frag_x;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];
  if (threadIdx.x < R) smax[threadIdx.x] = 0u;
  const int per_row = pitch - W;
  for (int i = threadIdx.x; i < R * per_row; i += blockDim.x) {
    const int r = i / per_row;
    const int c = i - r * per_row;
    const int col = r * pitch + (c < kx ? c : c + W);
    Ph[col] = 0;
    if (TERMS == 3) Pl[col] = 0;
  }

  const float M = key2f(b.max_keys[J.src_key]);
  const float negM = -M;
  const float e_fill = expf(kLogZero + negM);
  const int *oxo = shifts + J.shifts + 4 * R;
  const int *oyo = oxo + R;
  const bool zero_src = J.src_kind == kSrcZero;
  const UnaryMask mk = J.src_mask;
  const bool use_mask = J.src_kind == kSrcUnary && (mk.rot_bits != 0ull || mk.strip != 0);
  for (int x = threadIdx.x; x < W; x += blockDim.x) {
    float o[R];
    row_column<R>(cptr(b, J.src), oxo, oyo, taps + J.taps_r, negM, e_fill, H, W, H * W, J.delta, J.rot_mode,
                  J.rot_len, zero_src, use_mask, mk, y, x, o);
#pragma unroll
    for (int ro = 0; ro < R; ++ro) {
      unsigned short hi, lo;
      split_bf16(o[ro], hi, lo);
      Ph[ro * pitch + kx + x] = hi;
      if (TERMS == 3) Pl[ro * pitch + kx + x] = lo;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int m = min(lane & 31, R - 1);
  const int kq = (lane >> 5) << 3;
  const long long HW = (long long)H * W;
  float *T = mptr(b, J.tmp) + (long long)y * W;
  const int ntile = (W + 31) >> 5;
  float rmax[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) rmax[i] = 0.0f;
  for (int t = wave; t < ntile; t += nw) {
    const int x0 = t << 5;
    const unsigned short *ah = Ph + m * pitch + x0 + kq;
    const unsigned short *al = Pl + m * pitch + x0 + kq;
    f32x16_t acc = {};
    for (int kk = 0; kk < KK; ++kk) {
      const bf16x8_t a_hi = ld_frag(ah + 16 * kk);
      const bf16x8_t b_hi = ld_frag(fr + kk * 64 + lane);
      acc = mfma(a_hi, b_hi, acc);
      if (TERMS == 3) {
        const bf16x8_t a_lo = ld_frag(al + 16 * kk);
        const bf16x8_t b_lo = ld_frag(fr + (KK + kk) * 64 + lane);
        acc = mfma(a_lo, b_hi, acc);
        acc = mfma(a_hi, b_lo, acc);
      }
    }
    const int x = x0 + (lane & 31);
    if (x < W) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int r = 8 * (i >> 2) + ((lane >> 5) << 2) + (i & 3);
        if (r < R) {
          T[r * HW + x] = acc[i];
          rmax[i] = fmaxf(rmax[i], acc[i]);
        }
      }
    }
  }
  // slice maxima: the 32 lanes of each half hold the same rows
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    unsigned k = f2key(rmax[i]);
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) k = max(k, (unsigned)__shfl_xor((int)k, o, 64));
    const int r = 8 * (i >> 2) + ((lane >> 5) << 2) + (i & 3);
    if ((lane & 31) == 0 && r < R) atomicMax(smax + r, k);
  }
  __syncthreads();
  if (threadIdx.x < R && smax[threadIdx.x]) atomicMax(b.max_keys + J.slice_key + threadIdx.x, smax[threadIdx.x]);
}

// ---------------------------------------------------- column pass (y filter)
// One block per (message, slice r, strip of 64 columns), walking down the
// strip in tiles of 64 output rows.  The source rows a tile needs
// (32 + 16 KK: the wave's 32 rows plus the band) sit in an LDS ring, one
// column per 16-byte-aligned run of rows (the GEMM operand reads 8 rows of
// one column per lane); the next tile reuses all but 64 of them.  Each wave
// computes a 32 x 32 quarter of the tile, then the epilogue of the FP32
// column pass.
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
  const int nf = TERMS == 3 ? 2 : 1;
  const int nrows = mfma_ring_rows(KK);
  const int cs = mfma_ring_pitch(KK);
  uint4 *fr = lds;
  unsigned short *Rh = reinterpret_cast<unsigned short *>(lds + nf * KK * 64);
  unsigned short *Rl = Rh + 64 * cs;
  const uint4 *gf = frags + J.frag_y;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int *sh = shifts + J.shifts;
  const int ox = sh[2 * R + r];
  const int oy = sh[3 * R + r];
  const int ky = J.ky;
  const long long HW = (long long)H * W;
  const float *Tr = cptr(b, J.tmp) + r * HW;
  float scale, unscale;
  slice_scale(key2f(b.max_keys[J.slice_key + r]), scale, unscale);
  const float M = key2f(b.max_keys[J.src_key]);
  float *dst = mptr(b, J.dst);
  const float *dun = J.dst_unary.buf >= 0 ? cptr(b, J.dst_unary) : nullptr;
  float *dst2 = J.dst2.buf >= 0 ? mptr(b, J.dst2) : nullptr;
  const bool need_u = (J.dst_kind == kDstUpNode && J.dst_add_unary) || J.dst_kind == kDstDown;
  const bool need_b = J.dst_kind == kDstDown && J.dst_base == 0;

  // staging: this thread loads column lcol (= its LDS column lane) of T
  const int lcol = xt * 64 + lane + ox;
  const bool lcol_ok = xt * 64 + lane < W && lcol >= 0 && lcol < W;
  const float *Tc = Tr + (lcol_ok ? lcol : 0);
  auto src_row = [&](int row) -> float {
    const bool in = lcol_ok && row >= 0 && row < H;
    const float v = Tc[(long long)(in ? row : 0) * W];
    return in ? v * scale : 0.0f;
  };
  auto put = [&](int prow, float v) {
    unsigned short hi, lo;
    split_bf16(v, hi, lo);
    Rh[lane * cs + prow] = hi;
    if (TERMS == 3) Rl[lane * cs + prow] = lo;
  };
  // tile rows are source rows ty0 + oy - ky + [0, nrows), at ring row (base + row) mod nrows
  for (int rr = wave; rr < nrows; rr += 4) put(rr, src_row(oy - ky + rr));
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
  DstAcc da;
  int base = 0;
  for (int ty0 = 0; ty0 < H; ty0 += 64) {
    const bool more = ty0 + 64 < H;
    float pre[16];
    if (more) {
      const int s1 = ty0 + 64 + oy - ky + nrows - 64;
#pragma unroll
      for (int k = 0; k < 16; ++k) pre[k] = src_row(s1 + wave + 4 * k);
    }
    f32x16_t acc = {};
    for (int kk = 0; kk < KK; ++kk) {
      int lrow = base + ry + 16 * kk + kq;
      if (lrow >= nrows) lrow -= nrows;
      const bf16x8_t a_hi = ld_frag(fr + kk * 64 + lane);
      const bf16x8_t b_hi = ld_frag(bh_col + lrow);
      acc = mfma(a_hi, b_hi, acc);
      if (TERMS == 3) {
        const bf16x8_t a_lo = ld_frag(fr + (KK + kk) * 64 + lane);
        const bf16x8_t b_lo = ld_frag(bl_col + lrow);
        acc = mfma(a_hi, b_lo, acc);
        acc = mfma(a_lo, b_hi, acc);
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int y = ty0 + ry + 8 * (i >> 2) + ((lane >> 5) << 2) + (i & 3);
      if (!xin || y >= H) continue;
      const int ysrc = y + oy;
      const float msg = (col_ok && ysrc >= 0 && ysrc < H) ? log0(acc[i] * unscale) + M : kLogZero;
      const long long e = r * HW + (long long)y * W + x;
      const float pu = need_u ? dun[e] : 0.0f;
      const float pb = need_b ? dst[e] : 0.0f;
      dst_store(J, dst, dst2, e, xmask, msg, pu, pb, da);
    }
    if (more) {
      __syncthreads();  // every wave is done with this tile's ring rows
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        int prow = base + wave + 4 * k;
        if (prow >= nrows) prow -= nrows;
        put(prow, pre[k]);
      }
      base += 64;
      if (base >= nrows) base -= nrows;
      __syncthreads();
    }
  }
  dst_flush(b, J, dst2 != nullptr, da, lane);
}

// ------------------------------------------------------------ host side

size_t rowpass_mfma_lds_bytes(int R, int W, int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1;
  return 16 * (size_t)nf * KK * 64 + 2 * (size_t)nf * R * mfma_row_pitch(W, KK) + 4 * (size_t)R;
}
size_t colpass_mfma_lds_bytes(int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1;
  return 16 * (size_t)nf * KK * 64 + 2 * (size_t)nf * 64 * mfma_ring_pitch(KK);
}

hipError_t mfma_set_lds_limit(size_t row_bytes, size_t col_bytes) {
  auto set = [](const void *f, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  };
  hipError_t e = set(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 1>), row_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 3>), row_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_colpass_mfma<1>), col_bytes);
  if (e == hipSuccess) e = set(reinterpret_cast<const void *>(&k_colpass_mfma<3>), col_bytes);
  return e;
}

hipError_t launch_rowpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                               const uint4 *frags, int R, int terms, int max_h, int max_w, size_t lds_bytes,
                               hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  if (R != 24 || (terms != 1 && terms != 3)) return hipErrorInvalidValue;
  dim3 grid(max_h, n_jobs);
  const int nt = std::min(512, std::max(64, (max_w + 63) / 64 * 64));
  if (terms == 3)
    hipLaunchKernelGGL((k_rowpass_mfma<24, 3>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, frags);
  else
    hipLaunchKernelGGL((k_rowpass_mfma<24, 1>), grid, dim3(nt), lds_bytes, s, b, jobs, taps, shifts, frags);
  return hipGetLastError();
}

hipError_t launch_colpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const uint4 *frags, const int *shifts,
                               int R, int terms, int max_w, size_t lds_bytes, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  const int xtiles = (max_w + 63) / 64;
  dim3 grid(R * xtiles, n_jobs);
  if (terms == 3)
    hipLaunchKernelGGL(k_colpass_mfma<3>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  else
    hipLaunchKernelGGL(k_colpass_mfma<1>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  return hipGetLastError();
}

}  // namespace ps
