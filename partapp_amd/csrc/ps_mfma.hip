// Banded-Toeplitz variant of a message's filters on the gfx950 matrix cores.
// Same data flow as the FP32 row / column passes (the forward shift, exp, log,
// inverse shift and the caller's accumulation are unchanged); the two
// truncated Gaussian passes (multi_array_filter.hpp:277-288 along x, :304-317
// along y) become GEMMs against a Toeplitz matrix of the taps, and the
// circular rotation filter (objectdetect_findrot.cpp:377-420) a GEMM against
// its circulant matrix (row pass, below):
//
//   out[m][n] = sum_k A[m][k] B[k][n],  band entries f[k - m] (y pass, A) or
//                                       f[k - n] (x pass, B), zero elsewhere
//
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  A 32-wide output tile
// needs K = 32 + 2k inputs, rounded up to KK steps of 16.  TERMS selects the
// precision: 1 multiplies bf16(data) by bf16(taps); 2 splits the data into
// bf16 hi + lo parts (data to about 16 significant bits, taps in bf16); 3
// splits both and adds hi*hi + lo*hi + hi*lo.  With TERMS 3 the row pass
// stores FP32 and the bf16 column pass rescales every column of a slice by a
// power of two taken from its first rows (the y filter never mixes columns),
// so the bf16 operands keep their range; the scale is undone exactly after
// the GEMM.  With TERMS 1 and 2 the row pass stores fp16 scaled by a power of
// two per message and the column pass runs on the fp16 MFMA.  The exponential
// of the row pass and the logarithm of the column pass use the hardware
// instructions (a few ulp).
//
// Operand layouts of the 32x32x16 bf16 MFMA (lane l, t = 0..7):
//   A: A[l % 32][(l / 32) * 8 + t]   B: B[(l / 32) * 8 + t][l % 32]
//   C: acc[i] = C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32]
// Both Toeplitz operands therefore hold f[16 kk + 8 (l / 32) + t - l % 32]
// in lane l at step kk: one fragment table per filter, built on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "ps_device.hpp"
#include "ps_internal.hpp"

namespace ps {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

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
__device__ __forceinline__ f32x16_t mfma16(f16x8_t a, f16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f16x8_t ld_frag16(const uint4 *p) { return *reinterpret_cast<const f16x8_t *>(p); }
// eight floats rounded to fp16
__device__ __forceinline__ f16x8_t to_f16x8(const float (&v)[8]) {
  f16x8_t r;
#pragma unroll
  for (int t = 0; t < 8; ++t) r[t] = (_Float16)v[t];
  return r;
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

// element at a 32-bit byte offset from a (wave-uniform) base
__device__ __forceinline__ float &at(float *p, unsigned byte_off) {
  return *reinterpret_cast<float *>(reinterpret_cast<char *>(p) + byte_off);
}
__device__ __forceinline__ float at(const float *p, unsigned byte_off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + byte_off);
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
// One block per (message, run of `rows` output rows); wave w gathers columns
// 64 w .. 64 w + 63 and computes the two 32-wide output tiles over them.  For
// each row y: the lane's column of all R slices (forward shift, masks,
// exp(b - M), with M the source grid's maximum) is stored as bf16 (hi, and lo
// for TERMS >= 2) in LDS rows P[slice][kx + x]; then each wave runs the x
// filter as the KK-step GEMM P (R x K, rows clamped to R - 1 beyond R) times
// the Toeplitz fragments, and the rotation filter as a second GEMM F (R x R,
// circular taps) times that result, taken straight from the accumulator
// registers: lane l holds rows 8 (i / 4) + 4 (l / 32) + i % 4 of column l % 32
// in accumulator i, which is the B operand of k-step i / 8 once the K index
// is renumbered, so the host stores F with its columns permuted accordingly.
// Both filters are linear and act on different axes, so the order x filter
// then rotation equals the reference's rotation then x filter.  The rotation
// GEMM splits both its operands into bf16 hi + lo for every TERMS: its taps
// rounded to bf16 bias every message by up to 2^-9 in the same direction, which
// several chained messages turn into an error of ~1e-3 on small grids.
//
// The loads of row y + 1 are issued before row y's GEMMs and stay in flight
// through them; every load is unconditional (addresses clamped into the
// grid) and every slice's masking is branch-free.  Waves whose tiles lie past
// a job's width (the block is sized for the widest job of the launch) only
// gather and take part in the barriers.
//
// R = 24: one 32-row slice tile (slices 24..31 padding), a 2-step rotation
// GEMM.  R = 48: two slice tiles for the x filter, the rotation GEMM over 3
// k-steps (slices 0..31 from the first tile's accumulator, 32..47 from the
// second's) into two output tiles.
//
// R = 24 at most 72 VGPRs (7 waves per SIMD): five 320-thread blocks per CU
// for a 300-wide grid instead of four, at the cost of two spilled VGPRs.
template <int R>
struct RowTiles {
  static constexpr int MT = (R + 31) / 32;  // 32-slice tiles of the x filter and of the rotation output
  static constexpr int KS = (R + 15) / 16;  // k-steps of the rotation GEMM
  static constexpr int NRES = R == 24 ? 12 : 24;  // outputs per lane and column tile
};

template <int R, int TERMS, bool T16, bool ROWMAX>
__device__ __forceinline__ void rowpass_mfma_body(Bases b, const MsgJob *__restrict__ jobs,
                                                  const int *__restrict__ shifts, const uint4 *__restrict__ frags,
                                                  int rows) {
  static_assert(R == 24 || R == 48, "the MFMA row pass takes 24 or 48 rotation slices");
  constexpr int MT = RowTiles<R>::MT, KS = RowTiles<R>::KS, NRES = RowTiles<R>::NRES;
  extern __shared__ uint4 lds[];
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int H = J.H, W = J.W;
  const int y0 = (int)bx * rows;
  if (y0 >= H) return;
  const int y1 = min(H, y0 + rows);
  const int kx = J.kx, KK = J.kk_x;
  constexpr int nf = TERMS == 3 ? 2 : 1;  // x tap parts
  constexpr int nr = 2;                   // rotation tap parts (always hi + lo)
  const int pitch = mfma_row_pitch(W, kx, KK);
  uint4 *fr = lds;                 // [nf][KK][64] Toeplitz fragments
  uint4 *frr = fr + nf * KK * 64;  // [nr][MT][KS][64] rotation fragments
  unsigned short *Ph = reinterpret_cast<unsigned short *>(frr + nr * MT * KS * 64);
  unsigned short *Pl = Ph + R * pitch;
  unsigned *wkey = reinterpret_cast<unsigned *>(Ph + (TERMS >= 2 ? 2 : 1) * R * pitch);  // [waves]

  const uint4 *gf = frags + J.frag_x;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];
  // the rotation taps: bf16 hi / lo, or fp16 hi / lo for an fp16 output
  const uint4 *gr = frags + (T16 ? J.frag_r16 : J.frag_r);
  for (int i = threadIdx.x; i < nr * MT * KS * 64; i += blockDim.x) frr[i] = gr[i];
  // the halo columns stay zero for every row of the block
  const int per_row = pitch - W;
  for (int i = threadIdx.x; i < R * per_row; i += blockDim.x) {
    const int r = i / per_row;
    const int c = i - r * per_row;
    const int col = r * pitch + (c < kx ? c : c + W);
    Ph[col] = 0;
    if (TERMS >= 2) Pl[col] = 0;
  }

  // exp(b - M) = 2^(b log2 e - M log2 e)
  constexpr float kLog2e = 1.4426950408889634f;
  const float M = J.src_kind == kSrcGrid || !T16 ? key2f(b.max_keys[J.src_key]) : 0.0f;
  const float *src = cptr(b, J.src);
  const bool zero_src = J.src_kind == kSrcZero;
  const UnaryMask mk = J.src_mask;
  const bool use_mask = J.src_kind == kSrcUnary && (mk.rot_bits != 0ull || mk.strip != 0);
  int Ps = J.src.pitch;  // the source's row pitch (the caller's unaries: W; the plan's grids: padded)
  int HW = H * Ps;       // R * H * pitch < 2^31 is checked when the plan is built
  int delta = J.delta;
  // forward offsets by output slice (the source slice is ro - delta): x at
  // so, y at so + R, the source slice clamped into [0, R) at so + 3 R; read
  // through the scalar cache per row (the offset is opaque to the loop
  // optimiser so that they do not occupy scalar registers throughout)
  int so = J.shifts + 4 * R;
  // The lane's column is re-read through an empty asm every row, so the loop
  // optimiser cannot hoist the per-slice addresses derived from it (2 R of
  // them) out of the row loop and hold them in registers for its lifetime.
  auto lane_column = [] {
    int c = threadIdx.x;
    asm volatile("" : "+v"(c));
    return c;
  };

  // raw source values of row y (addresses clamped into the grid: every load
  // is unconditional; the row address on the scalar unit, the column a byte
  // offset from it) and their masking when consumed; both read the offsets
  // through the scalar cache and are branch-free per slice
  auto offsets = [&](int (&ox)[R], int (&oy)[R]) {
    asm volatile("" : "+s"(so));
#pragma unroll
    for (int ro = 0; ro < R; ++ro) {
      ox[ro] = shifts[so + ro];
      oy[ro] = shifts[so + R + ro];
    }
  };
  auto gather = [&](int y, float (&v)[R]) {
    asm volatile("" : "+s"(so), "+s"(HW), "+s"(Ps));
    const int x4 = lane_column() * 4, xmax4 = (W - 1) * 4;
#pragma unroll
    for (int ro = 0; ro < R; ++ro) {
      const int yy = min(max(y + shifts[so + R + ro], 0), H - 1);
      const char *srow = reinterpret_cast<const char *>(src) + (unsigned)(shifts[so + 3 * R + ro] * HW + yy * Ps) * 4u;
      const int xx4 = min(max(x4 + shifts[so + ro] * 4, 0), xmax4);
      v[ro] = *reinterpret_cast<const float *>(srow + (unsigned)xx4);
    }
  };
  // exp(b - M) = 2^(b log2 e - sub) of row y into the LDS rows, sub = M log2 e
  // (M the source grid's maximum) or, for an fp16 output, the row's exponent
  // less x16_shift (the values then carry that power of two)
  auto consume = [&](int y, const float (&v)[R], float sub) {
    int ox[R], oy[R];
    offsets(ox, oy);
    const int x = lane_column();
    float e[R];
    if (zero_src) {
      // a leaf without a detector keeps the zero log posterior it was
      // allocated with (no unary added, objectdetect_findrot.cpp:652-654):
      // exp(0 - M) = 1 (its max M is 0, and its row exponent 0), i.e.
      // 2^-sub, wherever the warp has a source pixel, exp(LOG_ZERO) = 0
      // elsewhere
      const float one = __builtin_amdgcn_exp2f(-sub);
#pragma unroll
      for (int ro = 0; ro < R; ++ro) {
        const bool ok = (unsigned)(y + oy[ro]) < (unsigned)H && (unsigned)(x + ox[ro]) < (unsigned)W;
        e[ro] = ok ? one : 0.0f;
      }
    } else if (use_mask) {
      // no source pixel, or a masked unary state: LOG_ZERO (objectdetect_findrot.cpp:342-365)
#pragma unroll
      for (int ro = 0; ro < R; ++ro) {
        const int xx = x + ox[ro];
        const int rs = ro - delta;
        const bool rot_masked = (unsigned)rs < (unsigned)R && ((mk.rot_bits >> (rs & 63)) & 1ull);
        const bool ok = (unsigned)(y + oy[ro]) < (unsigned)H && (unsigned)xx < (unsigned)W && !rot_masked &&
                        xx >= mk.strip && xx < W - mk.strip;
        e[ro] = __builtin_amdgcn_exp2f(fmaf(ok ? v[ro] : kLogZero, kLog2e, -sub));
      }
    } else {
#pragma unroll
      for (int ro = 0; ro < R; ++ro) {
        const int xx = x + ox[ro];
        const bool ok = (unsigned)(y + oy[ro]) < (unsigned)H && (unsigned)xx < (unsigned)W;
        e[ro] = __builtin_amdgcn_exp2f(fmaf(ok ? v[ro] : kLogZero, kLog2e, -sub));
      }
    }
    if (x < W) {
      // LDS addresses of slice ro's entries, advanced by one row pitch per
      // slice (an empty asm keeps the R addresses from being formed up front
      // and held in scalar registers)
      typedef __attribute__((address_space(3))) unsigned short lds_u16;
      lds_u16 *qh = (lds_u16 *)(Ph + kx + x), *ql = (lds_u16 *)(Pl + kx + x);
#pragma unroll
      for (int ro = 0; ro < R; ++ro) {
        const __bf16 hi = (__bf16)e[ro];
        *qh = __builtin_bit_cast(unsigned short, hi);
        qh += pitch;
        asm volatile("" : "+v"(qh));
        if (TERMS >= 2) {
          *ql = bf16_bits(e[ro] - (float)hi);
          ql += pitch;
          asm volatile("" : "+v"(ql));
        }
      }
    }
  };

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = min(lane & 31, R - 1);
  const int h = lane >> 5;
  const int kq = h << 3;
  const int ntile = (W + 31) >> 5;
  const int t0 = min(2 * wave, ntile - 1), t1 = min(2 * wave + 1, ntile - 1);
  // byte offsets of this lane's outputs: slice 8 (i / 4) + 4 h + i % 4, column
  // 32 t + lane % 32, from the row's start
  // T16: the output is stored as fp16 of value * 2^t16_shift (the shift
  // keeps the largest possible value near 2^14); the rows were exponentiated
  // with 2^x16_shift, which the store exchanges for it
  constexpr unsigned eb = T16 ? 2u : 4u;  // bytes per element
  const float t_scale = T16 ? ldexpf(1.0f, J.t16_shift - J.x16_shift) : 1.0f;
  const int tp = J.tpitch;
  unsigned slice_b = (unsigned)H * (unsigned)tp * eb;
  const unsigned ob0 = (unsigned)(4 * h) * slice_b + (unsigned)(32 * t0 + (lane & 31)) * eb;
  const unsigned ob1 = (unsigned)(4 * h) * slice_b + (unsigned)(32 * t1 + (lane & 31)) * eb;
  char *Tbase = reinterpret_cast<char *>(mptr(b, J.tmp));
  // fp16 rows take the first half of the output grid's bytes; their
  // exponents follow them
  int *rexp = reinterpret_cast<int *>(Tbase + (size_t)R * H * tp * 2);

  // x filter and rotation of one 32-column tile; res[i] is output slice
  // 8 (i / 4) + 4 h + i % 4 (< 24) of column 32 t + lane % 32 (R = 48: res[i]
  // for i < 16, then slice 32 + 8 (j / 4) + 4 h + j % 4 in res[16 + j], j < 8)
  auto tile = [&](int t, float (&res)[NRES]) {
    if constexpr (R == 48) {
      f32x16_t c[2];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        const int mrow = min(32 * mt + (lane & 31), R - 1);
        const unsigned short *ah = Ph + mrow * pitch + (t << 5) + kq;
        const unsigned short *al = Pl + mrow * pitch + (t << 5) + kq;
        f32x16_t cc = {};
        for (int kk = 0; kk < KK; ++kk) {
          const bf16x8_t a_hi = ld_frag(ah + 16 * kk);
          const bf16x8_t b_hi = ld_frag(fr + kk * 64 + lane);
          cc = mfma(a_hi, b_hi, cc);
          if (TERMS >= 2) cc = mfma(ld_frag(al + 16 * kk), b_hi, cc);
          if (TERMS == 3) cc = mfma(a_hi, ld_frag(fr + (KK + kk) * 64 + lane), cc);
        }
        c[mt] = cc;
      }
      // k-step ks takes accumulator tile ks / 2, registers 8 (ks % 2) ..
      f32x16_t d[2] = {};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        float w[8];
#pragma unroll
        for (int t8 = 0; t8 < 8; ++t8) w[t8] = c[ks >> 1][8 * (ks & 1) + t8];
        if (T16) {  // data as fp16, taps fp16 hi + lo (two products)
          const f16x8_t xb = to_f16x8(w);
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            d[mt] = mfma16(ld_frag16(frr + 64 * (mt * KS + ks) + lane), xb, d[mt]);
            d[mt] = mfma16(ld_frag16(frr + 64 * (MT * KS) + 64 * (mt * KS + ks) + lane), xb, d[mt]);
          }
        } else {
          bf16x8_t bh, bl;
          split8(w, bh, bl);
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            const bf16x8_t f = ld_frag(frr + 64 * (mt * KS + ks) + lane);
            d[mt] = mfma(f, bh, d[mt]);
            d[mt] = mfma(f, bl, d[mt]);
            d[mt] = mfma(ld_frag(frr + 64 * (MT * KS) + 64 * (mt * KS + ks) + lane), bh, d[mt]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) res[i] = d[0][i];
#pragma unroll
      for (int i = 0; i < 8; ++i) res[16 + i] = d[1][i];
      return;
    }
    const unsigned short *ah = Ph + m * pitch + (t << 5) + kq;
    const unsigned short *al = Pl + m * pitch + (t << 5) + kq;
    f32x16_t c = {};
    for (int kk = 0; kk < KK; ++kk) {
      const bf16x8_t a_hi = ld_frag(ah + 16 * kk);
      const bf16x8_t b_hi = ld_frag(fr + kk * 64 + lane);
      c = mfma(a_hi, b_hi, c);
      if (TERMS >= 2) c = mfma(ld_frag(al + 16 * kk), b_hi, c);
      if (TERMS == 3) c = mfma(a_hi, ld_frag(fr + (KK + kk) * 64 + lane), c);
    }
    // rotation: k-step s takes accumulators 8 s .. 8 s + 7.  An fp16 output:
    // data rounded to fp16 (the x filter's output at most ~2^14 there), taps
    // split fp16 hi + lo (two products); else data and taps split bf16 hi +
    // lo (three products: hi*hi, hi*lo, lo*hi)
    f32x16_t d = {};
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      float w[8];
#pragma unroll
      for (int t8 = 0; t8 < 8; ++t8) w[t8] = c[8 * st + t8];
      if (T16) {
        const f16x8_t xb = to_f16x8(w);
        d = mfma16(ld_frag16(frr + 64 * st + lane), xb, d);
        d = mfma16(ld_frag16(frr + 128 + 64 * st + lane), xb, d);
      } else {
        bf16x8_t bh, bl;
        split8(w, bh, bl);
        const bf16x8_t f = ld_frag(frr + 64 * st + lane);
        d = mfma(f, bh, d);
        d = mfma(f, bl, d);
        d = mfma(ld_frag(frr + 128 + 64 * st + lane), bh, d);
      }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) res[i] = d[i];
  };
  auto store_tile = [&](unsigned ob, char *Trow, const float (&res)[NRES]) {
#pragma unroll
    for (int k = 0; k < NRES; ++k) {
      const int i = k < 16 ? k : k - 16;
      // a 32-bit offset from the row's (scalar) base
      char *p = Trow + (ob + (unsigned)((k < 16 ? 0 : 32) + 8 * (i >> 2) + (i & 3)) * slice_b);
      if (T16)
        *reinterpret_cast<_Float16 *>(p) = (_Float16)(res[k] * t_scale);
      else
        *reinterpret_cast<float *>(p) = res[k];
    }
  };

  // T16: row y is stored as 2^(b log2 e - e_y) and e_y beside the output.
  // A source the sweep wrote itself (every source but a leaf's) has its
  // maximum M in the max keys: e_y = ceil(M log2 e) for every row.  A leaf's
  // unary (whose maximum no kernel computes) takes e_y = ceil(m log2 e), m
  // the largest value row y's loads read (each load reads a value of the
  // grid, masked or not: a bound of the row's unmasked values), from the
  // waves' maxima (order-preserving keys, one LDS slot per wave) published
  // before the barrier that ends the previous row; a leaf without a
  // detector has e_y = 0 (its values are 0 or LOG_ZERO).
  const bool row_max = ROWMAX && J.src_kind == kSrcUnary;
  const int e_grid = T16 && J.src_kind == kSrcGrid ? (int)ceilf(M * kLog2e) : 0;
  const int nw = (int)(blockDim.x >> 6);
  auto publish = [&](const float (&v)[R]) {
    float mx = v[0];
#pragma unroll
    for (int ro = 1; ro < R; ++ro) mx = fmaxf(mx, v[ro]);
    // wave maximum; the lane index is re-read through an empty asm so that
    // the six exchange addresses are not hoisted out of the row loop
    int self = lane_column() & 63;
    unsigned k = f2key(mx);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = max(k, (unsigned)__builtin_amdgcn_ds_bpermute((self ^ o) << 2, (int)k));
    if (lane == 0) wkey[wave] = k;
  };
  auto row_exponent = [&]() -> int {
    if (!row_max) return e_grid;
    unsigned k = 0u;
    for (int w = 0; w < nw; ++w) k = max(k, wkey[w]);
    return (int)ceilf(key2f(k) * kLog2e);
  };

  // Row y: its values into LDS, the loads of row y + 1 issued (in flight
  // through the GEMMs), then both tiles and their stores.
  float v[R];
  gather(y0, v);
  if (row_max) {
    publish(v);
    __syncthreads();
  }
  for (int y = y0; y < y1; ++y) {
    // per-slice scalars are recomputed every row rather than hoisted
    asm volatile("" : "+s"(so), "+s"(HW), "+s"(delta), "+s"(slice_b));
    if (T16) {
      const int ey = row_exponent();
      if (threadIdx.x == 0) rexp[y] = ey;
      consume(y, v, (float)(ey - J.x16_shift));
    } else {
      consume(y, v, M * kLog2e);
    }
    __builtin_amdgcn_sched_barrier(0);
    gather(min(y + 1, y1 - 1), v);
    __syncthreads();  // row y is in LDS; every wave has read its maxima
    char *Trow = Tbase + (long long)y * tp * eb;
    float res[NRES];
    if (2 * wave < ntile) {  // waves past the grid's width (a narrower job of the launch) idle
      tile(t0, res);
      store_tile(ob0, Trow, res);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (2 * wave + 1 < ntile) {
      tile(t1, res);
      store_tile(ob1, Trow, res);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (row_max && y + 1 < y1) publish(v);
    __syncthreads();  // P is rewritten by the next row
  }
}

// ROWMAX: the launch holds leaf messages, whose rows take their own exponents
template <int R, int TERMS, bool T16, bool ROWMAX>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(7))) void k_rowpass_mfma(
    Bases b, const MsgJob *__restrict__ jobs, const int *__restrict__ shifts, const uint4 *__restrict__ frags, int rows) {
  rowpass_mfma_body<R, TERMS, T16, ROWMAX>(b, jobs, shifts, frags, rows);
}
// 48 slices hold twice the accumulators: at most 128 VGPRs, 4 waves per SIMD
template <int TERMS, bool T16, bool ROWMAX>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_rowpass_mfma48(
    Bases b, const MsgJob *__restrict__ jobs, const int *__restrict__ shifts, const uint4 *__restrict__ frags, int rows) {
  rowpass_mfma_body<48, TERMS, T16, ROWMAX>(b, jobs, shifts, frags, rows);
}

// ---------------------------------------------------- column pass (y filter)
// One block per (message, slice r, strip of 64 columns), walking down the
// strip in tiles of 64 output rows.  The source rows a tile needs (32 + 16 KK:
// a wave's 32 rows plus the band), taken at the columns x + ox the inverse
// warp reads, scaled by the column's power of two and split into bf16 hi / lo,
// sit in an LDS ring with one column per run of rows (the GEMM operand reads 8
// rows of one column per lane; staging writes 16 rows of one column per
// thread); the next tile reuses all but 64 of them.  Each wave computes a
// 32 x 32 quarter of the tile, then the caller's accumulation.
//
// The accumulation is a template (EPI), so that every load of a tile is
// unconditional straight-line code: per tile a wave issues its epilogue
// operands, then the next tile's ring rows (always 16 loads, clamped into the
// grid), and the epilogue waits for its operands only while the ring loads
// stay in flight through it.
template <int TERMS, int EPI>
__device__ __forceinline__ void colpass_mfma_strip(const Bases &b, const MsgJob &J, const uint4 *__restrict__ frags,
                                                   const int *__restrict__ shifts, int R, int r, int xt,
                                                   uint4 *lds) {
  constexpr bool kNeedU = Epi<EPI>::kNeedU;
  constexpr bool kNeedB = Epi<EPI>::kNeedB;
  const int H = J.H, W = J.W;
  const int KK = J.kk_y;
  const int ky = J.ky;
  int tp = J.tpitch;
  const int nrows = mfma_ring_rows(KK);
  const int cs = mfma_ring_pitch(KK);
  constexpr int nf = TERMS == 3 ? 2 : 1;  // tap parts
  uint4 *fr = lds;                        // [nf][KK][64] Toeplitz fragments
  unsigned short *Rh = reinterpret_cast<unsigned short *>(lds + nf * KK * 64);
  unsigned short *Rl = Rh + 64 * cs;
  float *cexp = reinterpret_cast<float *>(Rl + (TERMS >= 2 ? 64 * cs : 0));  // [64] exponent of each column
  unsigned *cmax = reinterpret_cast<unsigned *>(cexp + 64);                  // [64] max of the first rows

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint4 *gf = frags + J.frag_y;
  for (int i = threadIdx.x; i < nf * KK * 64; i += blockDim.x) fr[i] = gf[i];
  if (threadIdx.x < 64) cmax[threadIdx.x] = 0u;
  const int ox = shifts[J.shifts + 2 * R + r];
  const int oy = shifts[J.shifts + 3 * R + r];
  const float M = key2f(b.max_keys[J.src_key]);
  // row pitches: pa of dst (and of the unary a marginal adds), pitch_b of
  // the grids at the other pitch (Epi::kSide)
  const int pa = J.dst.pitch, pitch_b = epi_pitch_b<EPI>(J);
  float *dst = mptr(b, J.dst) + r * H * pa;
  const float *dun = kNeedU ? cptr(b, J.dst_unary) + r * H * J.dst_unary.pitch : dst;
  float *dst2 = kNeedB ? mptr(b, J.dst2) + r * H * pitch_b : dst;
  const float *bel = kNeedB ? cptr(b, J.dst_bel) + r * H * pitch_b : dst;
  const bool leaf_u = J.dst_base == 1;  // kEpiDownLeaf: a detected part (else its base is 0)
  const UnaryMask dm = J.dst_mask;

  // staging: this thread loads column xt * 64 + lane + ox of the row-pass
  // output, 16 consecutive rows at a time.  The loads go to the nearest row
  // and column inside the grid; rows and columns outside it are zeroed where
  // the values are consumed (mask16), so no load sits under a branch.
  const int scol = xt * 64 + lane;
  const int lcol = scol + ox;
  const bool lcol_ok = scol < W && lcol >= 0 && lcol < W;
  const float *Tc = cptr(b, J.tmp) + (long long)r * H * tp + min(max(lcol, 0), W - 1);
  float scale = 1.0f;
  auto load16 = [&](int row0, float (&v)[16]) {  // row0 is wave-uniform
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = Tc[min(max(row0 + k, 0), H - 1) * tp];
  };
  auto mask16 = [&](int row0, float (&v)[16]) {
    if (!lcol_ok) {
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = 0.0f;
    }
    if (row0 < 0 || row0 + 16 > H) {
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = (row0 + k >= 0 && row0 + k < H) ? v[k] : 0.0f;
    }
  };
  // a 16-row chunk masked, scaled and split into bf16 hi / lo, then its LDS store
  auto split16 = [&](int row0, float (&v)[16], bf16x8_t &h0, bf16x8_t &l0, bf16x8_t &h1, bf16x8_t &l1) {
    mask16(row0, v);
    float w[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = v[k] * scale;
    split8(*reinterpret_cast<const float(*)[8]>(&w[0]), h0, l0);
    split8(*reinterpret_cast<const float(*)[8]>(&w[8]), h1, l1);
  };
  auto store16 = [&](int prow, const bf16x8_t &h0, const bf16x8_t &l0, const bf16x8_t &h1, const bf16x8_t &l1) {
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
    if (c < nrows / 16) load16(oy - ky + 16 * c, fill[q]);
  }
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) {
      mask16(oy - ky + 16 * c, fill[q]);
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
    if (wave == 0) cexp[lane] = log2f(unsc);  // an integer in [-60, 60]
  }
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) {
      bf16x8_t h0, l0, h1, l1;
      split16(oy - ky + 16 * c, fill[q], h0, l0, h1, l1);
      store16(16 * c, h0, l0, h1, l1);
    }
  }
  __syncthreads();

  const int cx = (wave & 1) * 32, ry = (wave >> 1) * 32;
  const int n = lane & 31, h = lane >> 5;
  const int kq = h << 3;
  const int x = xt * 64 + cx + n;
  const bool xin = x < W;
  const bool col_ok = xin && x + ox >= 0 && x + ox < W;
  const bool xmask = ((dm.rot_bits >> r) & 1ull) || x < dm.strip || x >= W - dm.strip;
  const unsigned short *bh_col = Rh + (cx + n) * cs;
  const unsigned short *bl_col = Rl + (cx + n) * cs;
  // log(v 2^e) + M = log2(v) ln 2 + (M + e ln 2)
  constexpr float kLn2 = 0.693147180559945309f;
  const float cM = fmaf(cexp[cx + n], kLn2, M);
  const int xo = xin ? x : W - 1;
  // the plan's own grids are written over their whole row pitch (LOG_ZERO in
  // the padding columns [W, pitch)): a 128-byte line written in parts by
  // different waves costs more than its bytes.  Loads keep the clamped column.
  const bool pad_a = !xin && x < pa;                           // a padding column of dst
  const bool pad_b = Epi<EPI>::kNeedB && !xin && x < pitch_b;  // of the next source
  const int xa = pad_a ? x : xo, xb = pad_b ? x : xo;
  unsigned Wb = (unsigned)pa * 4u, Wb2 = (unsigned)pitch_b * 4u;
  constexpr bool kUs = EPI == kEpiUpNodeU;  // the unary at the side pitch (a chain-node belief)

  Epi<EPI> acc;
  int base = 0;
  for (int ty0 = 0; ty0 < H; ty0 += 64) {
    // the per-row offsets are recomputed every tile rather than hoisted into
    // registers for the whole loop
    asm volatile("" : "+s"(Wb), "+s"(Wb2), "+s"(tp));
    // epilogue operands of this lane's 16 outputs, rows yl + 8 (i / 4) + i % 4,
    // at 32-bit byte offsets from the slice (rows past the grid's end, in the
    // last tile, read its last row)
    const int yl = ty0 + ry + 4 * h;
    const unsigned eb0 = (unsigned)(yl * pa + xa) * 4u;
    const unsigned ebs0 = Epi<EPI>::kSide ? (unsigned)(yl * pitch_b + xb) * 4u : eb0;
    float pu[16], pb[16];
    if (kNeedU || kNeedB) {
      if (ty0 + 64 <= H) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const unsigned dy = (unsigned)(8 * (i >> 2) + (i & 3));
          const unsigned eb = eb0 + dy * Wb, ebs = ebs0 + dy * Wb2;
          if (kNeedU) pu[i] = at(dun, kUs ? ebs : eb);
          if (kNeedB) pb[i] = at(bel, ebs);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int yc = min(yl + 8 * (i >> 2) + (i & 3), H - 1);
          const unsigned eb = (unsigned)(yc * pa + xa) * 4u, ebs = (unsigned)(yc * pitch_b + xb) * 4u;
          if (kNeedU) pu[i] = at(dun, kUs ? ebs : eb);
          if (kNeedB) pb[i] = at(bel, ebs);
        }
      }
    }
    // the next tile's new ring rows (in the last tile: clamped rows, unused)
    const bool more = ty0 + 64 < H;
    const int pre_row = ty0 + oy - ky + nrows + 16 * wave;
    float pre[16];
    load16(pre_row, pre);

    f32x16_t c = {};
    int lrow = base + ry + kq;
    if (lrow >= nrows) lrow -= nrows;
    for (int kk = 0; kk < KK; ++kk) {
      const bf16x8_t a_hi = ld_frag(fr + kk * 64 + lane);
      const bf16x8_t b_hi = ld_frag(bh_col + lrow);
      c = mfma(a_hi, b_hi, c);
      if (TERMS >= 2) c = mfma(a_hi, ld_frag(bl_col + lrow), c);
      if (TERMS == 3) c = mfma(ld_frag(fr + (KK + kk) * 64 + lane), b_hi, c);
      lrow += 16;
      if (lrow >= nrows) lrow -= nrows;
    }
    // the next tile's ring rows, split, before any store of this tile is
    // issued (that wait covers this tile's operands as well, all issued
    // before the GEMM)
    bf16x8_t nh0, nl0, nh1, nl1;
    if (more) split16(pre_row, pre, nh0, nl0, nh1, nl1);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int dy = 8 * (i >> 2) + (i & 3);
      const int y = yl + dy;
      const float lg = __builtin_amdgcn_logf(c[i]);  // -inf for 0
      const bool ok = col_ok && (unsigned)(y + oy) < (unsigned)H && lg > -INFINITY;
      const float msg = ok ? fmaf(lg, kLn2, cM) : kLogZero;
      if (xin && y < H)
        acc.store(dst, dst2, eb0 + (unsigned)dy * Wb, ebs0 + (unsigned)dy * Wb2, msg, kNeedU ? pu[i] : 0.0f,
                  kNeedB ? pb[i] : 0.0f, xmask, leaf_u);
      else if ((pad_a || pad_b) && y < H)
        acc.pad(dst, dst2, eb0 + (unsigned)dy * Wb, ebs0 + (unsigned)dy * Wb2, pad_a, pad_b);
    }
    if (more) {
      __syncthreads();  // every wave is done with this tile's ring rows
      int prow = base + 16 * wave;
      if (prow >= nrows) prow -= nrows;
      store16(prow, nh0, nl0, nh1, nl1);
      base += 64;
      if (base >= nrows) base -= nrows;
      __syncthreads();
    }
  }
  acc.flush(b, J, (long long)r * H * pa, lane);
}

// ---------------------------------------- column pass on an fp16 row-pass output
// The methods whose row pass stores fp16 run the y filter on the fp16 MFMA
// (v_mfma_f32_32x32x16_f16): the data operand is the stored value itself
// (exact, no split, no per-column rescale: fp16 values convert to FP32
// exactly), the taps are split into fp16 hi + lo, two products per k-step.
// The LDS ring holds one fp16 plane, half the bytes of the bf16 hi / lo pair,
// and the taps sit in LDS as the 8 shifted copies of f16_tap_copy_len (5 KB
// at KK = 7 where per-lane fragments take 14 KB), so that LDS no longer
// limits a CU to 4 blocks.  Block shape, staging order and epilogue as in
// colpass_mfma_strip.
// PRUNE: a pruned execution; a 64 x 64 destination tile whose keep flag is 0
// skips its GEMM and epilogue operands and is written as LOG_ZERO (zero
// probability: the pruned model sets those states' unaries to -infinity).
// The ring keeps staging every tile, so the next kept tile finds its rows.
template <int EPI, bool PRUNE, bool RS>
__device__ __forceinline__ void colpass_f16_strip(const Bases &b, const MsgJob &J, const uint4 *__restrict__ frags,
                                                  const int *__restrict__ shifts, int R, int r, int xt, uint4 *lds) {
  constexpr bool kNeedU = Epi<EPI>::kNeedU;
  constexpr bool kNeedB = Epi<EPI>::kNeedB;
  const int H = J.H, W = J.W;
  const int KK = J.kk_y;
  const int ky = J.ky;
  int tp = J.tpitch;
  const int nrows = mfma_ring_rows(KK);
  const int cs = mfma_ring_pitch(KK);
  const int q8 = f16_tap_copy_len(KK) / 8;  // 16-byte units per tap copy
  uint4 *fr = lds;  // [2][8][q8] fp16 tap copies, hi then lo
  unsigned short *Rg = reinterpret_cast<unsigned short *>(lds + 16 * q8);  // [64][cs] fp16 ring

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint4 *gf = frags + J.frag_y16;
  for (int i = threadIdx.x; i < 16 * q8; i += blockDim.x) fr[i] = gf[i];
  const int ox = shifts[J.shifts + 2 * R + r];
  const int oy = shifts[J.shifts + 3 * R + r];
  // the row pass stored row y as 2^(v log2 e - e_y) (times 2^t16_shift), e_y
  // beside the output.  RS (a leaf message in the launch: rows with their own
  // exponents): the rows are brought to the message's largest exponent emax
  // as they are staged, by the exact power-of-two scales 2^(e_y - emax) (0
  // below 2^-24), kept in LDS as fp16 pairs: sp[1 + i] scales rows i and
  // i + 1 (clamped into the grid; i = -1 .. H - 1), the pairs the ring
  // stores.  Without RS every row of a message has the same exponent.
  const int *rexp = reinterpret_cast<const int *>(reinterpret_cast<const char *>(cptr(b, J.tmp)) +
                                                  (size_t)R * H * tp * 2);
  int emax;
  unsigned *sp = nullptr;
  if (RS && J.src_kind == kSrcUnary) {
    int *xk = reinterpret_cast<int *>(Rg + 64 * cs);  // [4] the waves' largest row exponents
    sp = reinterpret_cast<unsigned *>(xk + 4);         // [H + 1] fp16 scale pairs
    int em = INT_MIN;
    for (int i = threadIdx.x; i < H; i += blockDim.x) em = max(em, rexp[i]);
    const unsigned k = wave_max_u32((unsigned)em ^ 0x80000000u);
    if (lane == 0) xk[wave] = (int)(k ^ 0x80000000u);
    __syncthreads();
    emax = max(max(xk[0], xk[1]), max(xk[2], xk[3]));
    auto scale_bits = [&](int row) -> unsigned {
      const int d = rexp[min(max(row, 0), H - 1)] - emax;
      return d >= -14 ? (unsigned)(d + 15) << 10 : (d >= -24 ? 1u << (d + 24) : 0u);
    };
    for (int i = (int)threadIdx.x - 1; i < H; i += blockDim.x) sp[1 + i] = scale_bits(i) | scale_bits(i + 1) << 16;
    __syncthreads();
  } else {
    emax = rexp[0];
  }
  // row pitches: pa of dst (and of the unary a marginal adds), pitch_b of
  // the grids at the other pitch (Epi::kSide)
  const int pa = J.dst.pitch, pitch_b = epi_pitch_b<EPI>(J);
  float *dst = mptr(b, J.dst) + r * H * pa;
  const float *dun = kNeedU ? cptr(b, J.dst_unary) + r * H * J.dst_unary.pitch : dst;
  float *dst2 = kNeedB ? mptr(b, J.dst2) + r * H * pitch_b : dst;
  const float *bel = kNeedB ? cptr(b, J.dst_bel) + r * H * pitch_b : dst;
  const bool leaf_u = J.dst_base == 1;
  const UnaryMask dm = J.dst_mask;

  // staging: column xt * 64 + lane + ox, 16 rows at a time (loads clamped into
  // the grid, rows / columns outside it zeroed before they reach LDS)
  const int scol = xt * 64 + lane;
  const int lcol = scol + ox;
  const bool lcol_ok = scol < W && lcol >= 0 && lcol < W;
  // rows in pairs as fp16 vectors: moved and masked as values (selects, no
  // arithmetic), which holds fewer registers than 16 zero-extended halves
  const _Float16 *Th =
      reinterpret_cast<const _Float16 *>(cptr(b, J.tmp)) + (long long)r * H * tp + min(max(lcol, 0), W - 1);
  auto load16 = [&](int row0, h2_t (&v)[8]) {  // row0 is wave-uniform
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      h2_t t;
      t.x = Th[min(max(row0 + 2 * k, 0), H - 1) * tp];
      t.y = Th[min(max(row0 + 2 * k + 1, 0), H - 1) * tp];
      v[k] = t;
    }
  };
  auto put16 = [&](int row0, int prow, const h2_t (&v)[8]) {
    const bool all = lcol_ok && row0 >= 0 && row0 + 16 <= H;
    h2_t *dstp = reinterpret_cast<h2_t *>(Rg + lane * cs + prow);
    unsigned s8[8];  // the scales, read before any ring store (which could alias them)
    if (RS && sp) {
#pragma unroll
      for (int k = 0; k < 8; ++k) s8[k] = sp[1 + min(max(row0 + 2 * k, -1), H - 1)];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool o0 = all || (lcol_ok && row0 + 2 * k >= 0 && row0 + 2 * k < H);
      const bool o1 = all || (lcol_ok && row0 + 2 * k + 1 >= 0 && row0 + 2 * k + 1 < H);
      h2_t m = RS && sp ? v[k] * __builtin_bit_cast(h2_t, s8[k]) : v[k];
      m.x = o0 ? m.x : (_Float16)0.0f;
      m.y = o1 ? m.y : (_Float16)0.0f;
      dstp[k] = m;
    }
  };
  // tile rows are source rows ty0 + oy - ky + [0, nrows), at ring row (base +
  // row) mod nrows
  constexpr int kFill = 3;  // chunks of 16 rows per thread in the first fill (nrows <= 160)
  h2_t fill[kFill][8];
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) load16(oy - ky + 16 * c, fill[q]);
  }
#pragma unroll
  for (int q = 0; q < kFill; ++q) {
    const int c = wave + 4 * q;
    if (c < nrows / 16) put16(oy - ky + 16 * c, 16 * c, fill[q]);
  }
  __syncthreads();

  const int cx = (wave & 1) * 32, ry = (wave >> 1) * 32;
  const int n = lane & 31, h = lane >> 5;
  const int kq = h << 3;
  const int x = xt * 64 + cx + n;
  const bool xin = x < W;
  const bool col_ok = xin && x + ox >= 0 && x + ox < W;
  const bool xmask = ((dm.rot_bits >> r) & 1ull) || x < dm.strip || x >= W - dm.strip;
  const unsigned short *b_col = Rg + (cx + n) * cs;
  // A fragment of k-step kk: taps 16 kk + c + t - 32 (t = 0..7), c = 8 h + 32 - n,
  // read from the copy shifted by c mod 8 at the aligned element 16 kk + 8 (c / 8)
  const int ca = 8 * h + 32 - n;
  const uint4 *fa = fr + (ca & 7) * q8 + (ca >> 3);
  // log(v 2^-s) + emax ln 2 = log2(v) ln 2 + (emax - s) ln 2, s the row
  // pass's fp16 shift
  constexpr float kLn2 = 0.693147180559945309f;
  const float cM = (float)(emax - J.t16_shift) * kLn2;
  const int xo = xin ? x : W - 1;
  // the plan's own grids are written over their whole row pitch (LOG_ZERO in
  // the padding columns [W, pitch)): a 128-byte line written in parts by
  // different waves costs more than its bytes.  Loads keep the clamped column.
  const bool pad_a = !xin && x < pa;                           // a padding column of dst
  const bool pad_b = Epi<EPI>::kNeedB && !xin && x < pitch_b;  // of the next source
  const int xa = pad_a ? x : xo, xb = pad_b ? x : xo;
  unsigned Wb = (unsigned)pa * 4u, Wb2 = (unsigned)pitch_b * 4u;
  constexpr bool kUs = EPI == kEpiUpNodeU;  // the unary at the side pitch (a chain-node belief)

  Epi<EPI> acc;
  int base = 0;
  const unsigned char *kp = PRUNE ? b.keep + J.keep_off + (long long)r * J.keep_ty * J.keep_tx + xt : nullptr;
  for (int ty0 = 0; ty0 < H; ty0 += 64) {
    asm volatile("" : "+s"(Wb), "+s"(Wb2), "+s"(tp));  // per-row offsets recomputed per tile, not hoisted
    const int yl = ty0 + ry + 4 * h;
    const unsigned eb0 = (unsigned)(yl * pa + xa) * 4u;
    const unsigned ebs0 = Epi<EPI>::kSide ? (unsigned)(yl * pitch_b + xb) * 4u : eb0;
    const bool kept = !PRUNE || kp[(ty0 >> 6) * J.keep_tx] != 0;
    float pu[16], pb[16];
    if ((kNeedU || kNeedB) && kept) {
      if (ty0 + 64 <= H) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const unsigned dy = (unsigned)(8 * (i >> 2) + (i & 3));
          const unsigned eb = eb0 + dy * Wb, ebs = ebs0 + dy * Wb2;
          if (kNeedU) pu[i] = at(dun, kUs ? ebs : eb);
          if (kNeedB) pb[i] = at(bel, ebs);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int yc = min(yl + 8 * (i >> 2) + (i & 3), H - 1);
          const unsigned eb = (unsigned)(yc * pa + xa) * 4u, ebs = (unsigned)(yc * pitch_b + xb) * 4u;
          if (kNeedU) pu[i] = at(dun, kUs ? ebs : eb);
          if (kNeedB) pb[i] = at(bel, ebs);
        }
      }
    }
    const bool more = ty0 + 64 < H;
    const int pre_row = ty0 + oy - ky + nrows + 16 * wave;
    h2_t pre[8];
    load16(pre_row, pre);

    if (kept) {
      f32x16_t c = {};
      int lrow = base + ry + kq;
      if (lrow >= nrows) lrow -= nrows;
      for (int kk = 0; kk < KK; ++kk) {
        const f16x8_t bv = *reinterpret_cast<const f16x8_t *>(b_col + lrow);
        c = mfma16(*reinterpret_cast<const f16x8_t *>(fa + 2 * kk), bv, c);
        c = mfma16(*reinterpret_cast<const f16x8_t *>(fa + 8 * q8 + 2 * kk), bv, c);
        lrow += 16;
        if (lrow >= nrows) lrow -= nrows;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int dy = 8 * (i >> 2) + (i & 3);
        const int y = yl + dy;
        const float lg = __builtin_amdgcn_logf(c[i]);  // -inf for 0
        const bool ok = col_ok && (unsigned)(y + oy) < (unsigned)H && lg > -INFINITY;
        const float msg = ok ? fmaf(lg, kLn2, cM) : kLogZero;
        if (xin && y < H)
          acc.store(dst, dst2, eb0 + (unsigned)dy * Wb, ebs0 + (unsigned)dy * Wb2, msg, kNeedU ? pu[i] : 0.0f,
                    kNeedB ? pb[i] : 0.0f, xmask, leaf_u);
        else if ((pad_a || pad_b) && y < H)
          acc.pad(dst, dst2, eb0 + (unsigned)dy * Wb, ebs0 + (unsigned)dy * Wb2, pad_a, pad_b);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int dy = 8 * (i >> 2) + (i & 3);
        if (yl + dy < H) {
          if (xin || pad_a) at(dst, eb0 + (unsigned)dy * Wb) = kLogZero;
          if (kNeedB && (xin || pad_b)) at(dst2, ebs0 + (unsigned)dy * Wb2) = kLogZero;
        }
      }
    }
    if (more) {
      __syncthreads();  // every wave is done with this tile's ring rows
      int prow = base + 16 * wave;
      if (prow >= nrows) prow -= nrows;
      put16(pre_row, prow, pre);
      base += 64;
      if (base >= nrows) base -= nrows;
      __syncthreads();
    }
  }
  acc.flush(b, J, (long long)r * H * pa, lane);
}

// HI: the highest epilogue kind among the launch's jobs.  The kernel holds
// only the kinds up to HI, so a launch without the five-grid downward
// epilogue is not held to its register count (UpRoot 86, UpNodeU 94,
// DownLeaf 98, DownMid 116 VGPRs + AGPRs).
template <int HI, bool PRUNE, bool RS>
__global__ __launch_bounds__(256) void k_colpass_f16(Bases b, const MsgJob *__restrict__ jobs,
                                                     const uint4 *__restrict__ frags, const int *__restrict__ shifts,
                                                     int R, int xtiles) {
  extern __shared__ uint4 lds[];
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int r = (int)bx / xtiles;
  const int xt = (int)bx - r * xtiles;
  if (xt * 64 >= J.W) return;
  const int kind = epi_kind(J);
  if (HI == kEpiUpRoot || kind == kEpiUpRoot)
    colpass_f16_strip<kEpiUpRoot, PRUNE, RS>(b, J, frags, shifts, R, r, xt, lds);
  else if (HI == kEpiUpNode || kind == kEpiUpNode)
    colpass_f16_strip<kEpiUpNode, PRUNE, RS>(b, J, frags, shifts, R, r, xt, lds);
  else if (HI == kEpiUpNodeU || kind == kEpiUpNodeU)
    colpass_f16_strip<kEpiUpNodeU, PRUNE, RS>(b, J, frags, shifts, R, r, xt, lds);
  else if (HI == kEpiDownLeaf || kind == kEpiDownLeaf)
    colpass_f16_strip<kEpiDownLeaf, PRUNE, RS>(b, J, frags, shifts, R, r, xt, lds);
  else
    colpass_f16_strip<kEpiDownMid, PRUNE, RS>(b, J, frags, shifts, R, r, xt, lds);
}

template <int TERMS>
__global__ __launch_bounds__(256) void k_colpass_mfma(Bases b, const MsgJob *__restrict__ jobs,
                                                      const uint4 *__restrict__ frags, const int *__restrict__ shifts,
                                                      int R, int xtiles) {
  extern __shared__ uint4 lds[];
  unsigned bx, by;
  xcd_block(bx, by);
  const MsgJob &J = jobs[by];
  if (skip_inst(b, J.inst)) return;
  const int r = (int)bx / xtiles;
  const int xt = (int)bx - r * xtiles;
  if (xt * 64 >= J.W) return;
  if (J.dst_kind == kDstUpRoot)
    colpass_mfma_strip<TERMS, kEpiUpRoot>(b, J, frags, shifts, R, r, xt, lds);
  else if (J.dst_kind == kDstUpNode && !J.dst_add_unary)
    colpass_mfma_strip<TERMS, kEpiUpNode>(b, J, frags, shifts, R, r, xt, lds);
  else if (J.dst_kind == kDstUpNode)
    colpass_mfma_strip<TERMS, kEpiUpNodeU>(b, J, frags, shifts, R, r, xt, lds);
  else if (J.dst_base != 0)
    colpass_mfma_strip<TERMS, kEpiDownLeaf>(b, J, frags, shifts, R, r, xt, lds);
  else
    colpass_mfma_strip<TERMS, kEpiDownMid>(b, J, frags, shifts, R, r, xt, lds);
}

// ------------------------------------------------------------ host side

size_t rowpass_mfma_lds_bytes(int R, int W, int kx, int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1, nd = terms >= 2 ? 2 : 1;
  const int rot = 2 * ((R + 31) / 32) * ((R + 15) / 16);  // hi / lo rotation fragments per lane
  return 16 * ((size_t)nf * KK + rot) * 64 + 2 * (size_t)nd * R * mfma_row_pitch(W, kx, KK) + 64;
}
size_t colpass_mfma_lds_bytes(int KK, int terms) {
  const int nf = terms == 3 ? 2 : 1, nd = terms >= 2 ? 2 : 1;
  return 16 * (size_t)nf * KK * 64 + 2 * (size_t)nd * 64 * mfma_ring_pitch(KK) + 8 * 64;
}

size_t colpass_f16_lds_bytes(int KK, int H) {
  return 2 * 8 * 2 * (size_t)f16_tap_copy_len(KK) + 2 * (size_t)64 * mfma_ring_pitch(KK) + 16 + 4 * ((size_t)H + 1);
}

hipError_t raise_lds_limit(const void *f, size_t bytes);

template <bool PRUNE, bool RS>
static hipError_t colpass_f16_set_lds_limit(size_t bytes) {
  hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_f16<kEpiUpRoot, PRUNE, RS>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_f16<kEpiUpNode, PRUNE, RS>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_f16<kEpiUpNodeU, PRUNE, RS>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_f16<kEpiDownLeaf, PRUNE, RS>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_f16<kEpiDownMid, PRUNE, RS>), bytes);
  return e;
}

template <bool ROWMAX>
static hipError_t rowpass_mfma_set_lds_limit(size_t bytes) {
  hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 1, true, ROWMAX>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 2, true, ROWMAX>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma48<1, true, ROWMAX>), bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma48<2, true, ROWMAX>), bytes);
  return e;
}

hipError_t mfma_set_lds_limit(size_t row_bytes, size_t col_bytes) {
  hipError_t e = rowpass_mfma_set_lds_limit<false>(row_bytes);
  if (e == hipSuccess) e = rowpass_mfma_set_lds_limit<true>(row_bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma<24, 3, false, false>), row_bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_rowpass_mfma48<3, false, false>), row_bytes);
  if (e == hipSuccess) e = raise_lds_limit(reinterpret_cast<const void *>(&k_colpass_mfma<3>), col_bytes);
  if (e == hipSuccess) e = colpass_f16_set_lds_limit<false, false>(col_bytes);
  if (e == hipSuccess) e = colpass_f16_set_lds_limit<false, true>(col_bytes);
  if (e == hipSuccess) e = colpass_f16_set_lds_limit<true, false>(col_bytes);
  if (e == hipSuccess) e = colpass_f16_set_lds_limit<true, true>(col_bytes);
  return e;
}

template <int R, bool ROWMAX>
static void launch_rowpass_t16(int terms, dim3 grid, int nt, size_t lds_bytes, hipStream_t s, const Bases &b,
                               const MsgJob *jobs, const int *shifts, const uint4 *frags, int rows) {
  if (R == 48) {
    if (terms == 2)
      hipLaunchKernelGGL((k_rowpass_mfma48<2, true, ROWMAX>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
    else
      hipLaunchKernelGGL((k_rowpass_mfma48<1, true, ROWMAX>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
  } else {
    if (terms == 2)
      hipLaunchKernelGGL((k_rowpass_mfma<24, 2, true, ROWMAX>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
    else
      hipLaunchKernelGGL((k_rowpass_mfma<24, 1, true, ROWMAX>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
  }
}

// leaf: the jobs include a leaf's message (a source without a grid maximum)
hipError_t launch_rowpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts,
                               const uint4 *frags, int R, int terms, int rows, int max_h, int max_w,
                               size_t lds_bytes, bool leaf, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  // one lane per column: the plan only routes grids of width <= 512 here
  if ((R != 24 && R != 48) || terms < 1 || terms > 3 || max_w > 512 || rows < 1) return hipErrorInvalidValue;
  dim3 grid((max_h + rows - 1) / rows, n_jobs);
  const int nt = std::min(512, std::max(64, (max_w + 63) / 64 * 64));
  if (terms == 3) {
    if (R == 48)
      hipLaunchKernelGGL((k_rowpass_mfma48<3, false, false>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
    else
      hipLaunchKernelGGL((k_rowpass_mfma<24, 3, false, false>), grid, dim3(nt), lds_bytes, s, b, jobs, shifts, frags, rows);
  } else if (R == 48) {
    if (leaf)
      launch_rowpass_t16<48, true>(terms, grid, nt, lds_bytes, s, b, jobs, shifts, frags, rows);
    else
      launch_rowpass_t16<48, false>(terms, grid, nt, lds_bytes, s, b, jobs, shifts, frags, rows);
  } else {
    if (leaf)
      launch_rowpass_t16<24, true>(terms, grid, nt, lds_bytes, s, b, jobs, shifts, frags, rows);
    else
      launch_rowpass_t16<24, false>(terms, grid, nt, lds_bytes, s, b, jobs, shifts, frags, rows);
  }
  return hipGetLastError();
}

template <bool PRUNE, bool RS>
static void launch_colpass_f16(int epi_hi, dim3 grid, size_t lds_bytes, hipStream_t s, const Bases &b,
                               const MsgJob *jobs, const uint4 *frags, const int *shifts, int R, int xtiles) {
  const dim3 blk(256);
  switch (epi_hi) {
    case kEpiUpRoot: hipLaunchKernelGGL((k_colpass_f16<kEpiUpRoot, PRUNE, RS>), grid, blk, lds_bytes, s, b, jobs, frags, shifts, R, xtiles); break;
    case kEpiUpNode: hipLaunchKernelGGL((k_colpass_f16<kEpiUpNode, PRUNE, RS>), grid, blk, lds_bytes, s, b, jobs, frags, shifts, R, xtiles); break;
    case kEpiUpNodeU: hipLaunchKernelGGL((k_colpass_f16<kEpiUpNodeU, PRUNE, RS>), grid, blk, lds_bytes, s, b, jobs, frags, shifts, R, xtiles); break;
    case kEpiDownLeaf: hipLaunchKernelGGL((k_colpass_f16<kEpiDownLeaf, PRUNE, RS>), grid, blk, lds_bytes, s, b, jobs, frags, shifts, R, xtiles); break;
    default: hipLaunchKernelGGL((k_colpass_f16<kEpiDownMid, PRUNE, RS>), grid, blk, lds_bytes, s, b, jobs, frags, shifts, R, xtiles); break;
  }
}

// t16: the row pass stored its output as fp16 (row-pass terms 1 and 2);
// epi_hi: the highest epilogue kind (EpiKind) among the jobs; leaf: the jobs
// include a leaf's message (fp16 rows with their own exponents)
hipError_t launch_colpass_mfma(const Bases &b, const MsgJob *jobs, int n_jobs, const uint4 *frags, const int *shifts,
                               int R, int terms, bool t16, int epi_hi, int max_w, size_t lds_bytes, bool leaf,
                               hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  if (terms < 1 || terms > 3 || (!t16 && terms != 3) || epi_hi < kEpiUpRoot || epi_hi > kEpiDownMid)
    return hipErrorInvalidValue;
  const int xtiles = (max_w + 63) / 64;
  dim3 grid(R * xtiles, n_jobs);
  if (t16) {  // fp16 operands: exact data, split taps, whatever the method's terms
    if (b.keep) {  // a pruned execution
      if (leaf)
        launch_colpass_f16<true, true>(epi_hi, grid, lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
      else
        launch_colpass_f16<true, false>(epi_hi, grid, lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
    } else if (leaf) {
      launch_colpass_f16<false, true>(epi_hi, grid, lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
    } else {
      launch_colpass_f16<false, false>(epi_hi, grid, lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
    }
  } else {  // an FP32 row-pass output (3-term row pass): bf16 hi / lo split of data and taps
    hipLaunchKernelGGL(k_colpass_mfma<3>, grid, dim3(256), lds_bytes, s, b, jobs, frags, shifts, R, xtiles);
  }
  return hipGetLastError();
}

}  // namespace ps
