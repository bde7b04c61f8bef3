// The first half of a message's row pass, shared by the FP32 and the MFMA
// row passes: for one output row y and column x, the forward-warped child
// belief of every rotation slice (one integer shift per slice,
// objectdetect_findrot.cpp:342-359), the upright / border masks of a leaf's
// unary, exp(b - M) (:362-365) and the circular rotation filter (:377-420).
#pragma once

#include <hip/hip_runtime.h>

#include "ps_device.hpp"
#include "ps_internal.hpp"

namespace ps {

// Circular rotation filter with LR taps on a register column
// (objectdetect_findrot.cpp:377-420, multi_array_filter.hpp:115-155):
// out[r] = sum_j in[(r - n + j) mod R] f[j], taps in ascending order.
template <int R, int LR>
__device__ __forceinline__ void rot_filter_fixed(const float (&e)[R], float (&o)[R], const float *fr) {
  constexpr int n = (LR - 1) / 2;
  float f[LR];
#pragma unroll
  for (int j = 0; j < LR; ++j) f[j] = fr[j];
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < LR; ++j) acc = fmaf(e[((ro - n + j) % R + R) % R], f[j], acc);
    o[ro] = acc;
  }
}
template <int R, int LR>
struct RotDispatch {
  static __device__ __forceinline__ void run(int lr, const float (&e)[R], float (&o)[R], const float *fr) {
    if (lr == LR) rot_filter_fixed<R, LR>(e, o, fr);
    else RotDispatch<R, LR - 2>::run(lr, e, o, fr);
  }
};
template <int R>
struct RotDispatch<R, -1> {
  static __device__ __forceinline__ void run(int, const float (&)[R], float (&o)[R], const float *) {
#pragma unroll
    for (int r = 0; r < R; ++r) o[r] = 0.0f;
  }
};

// Rotation-filtered exp(b - M) of column x of output row y for all R output
// slices.  Per OUTPUT slice ro the source slice is ro - delta, read at
// (y + oyo[ro], x + oxo[ro]); slices without a source carry a row offset
// that fails the bounds test.  The loads of all slices are issued before any
// of them is used.
template <int R>
__device__ __forceinline__ void row_column(const float *src, const int *oxo, const int *oyo, const float *fr,
                                           float negM, float e_fill, int H, int W, int HW32, int delta,
                                           int rot_mode, int rot_len, bool zero_src, bool use_mask,
                                           UnaryMask mk, int y, int x, float (&o)[R]) {
  // off < 0: no source pixel, the warp's LOG_ZERO fill
  int off[R];
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    const int yy = y + oyo[ro];
    const int xx = x + oxo[ro];
    const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
    off[ro] = ok ? (ro - delta) * HW32 + yy * W + xx : -1;
  }
  float e[R];
  if (zero_src) {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) e[ro] = 0.0f;
  } else {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) e[ro] = src[off[ro] < 0 ? 0 : off[ro]];
  }
  // masks, then exp(b - M) (:362-365)
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    float vv = e[ro];
    if (use_mask && masked(mk, ro - delta, x + oxo[ro], W)) vv = kLogZero;
    e[ro] = off[ro] < 0 ? e_fill : expf(vv + negM);
  }
  if (rot_mode == kRotFilter) {
    RotDispatch<R, (R % 2 == 1) ? R - 2 : R - 1>::run(rot_len, e, o, fr);
  } else if (rot_mode == kRotCopy) {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) o[ro] = e[ro];
  } else {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) o[ro] = 0.0f;
  }
}

// The same in two halves, so a block can issue the loads of its next row
// before it processes the current one.  row_loads fills v with the raw
// source values of column x of row y and returns bit ro set when output slice
// ro has a source pixel; row_finish applies the masks, exp(b - M) and the
// rotation filter.  FAST selects the hardware exponential (v_exp_f32 of
// x * log2(e), a few ulp) instead of libm's expf: for the BF16 passes, whose
// operands are rounded to 8 bits right after.
template <int R>
__device__ __forceinline__ unsigned long long row_loads(const float *src, const int *oxo, const int *oyo, int H,
                                                        int W, int HW32, int delta, bool zero_src, int y, int x,
                                                        float (&v)[R]) {
  static_assert(R <= 64, "one bit per slice");
  unsigned long long ok_bits = 0ull;
  int off[R];
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    const int yy = y + oyo[ro];
    const int xx = x + oxo[ro];
    const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
    off[ro] = ok ? (ro - delta) * HW32 + yy * W + xx : 0;
    ok_bits |= ok ? (1ull << ro) : 0ull;
  }
#pragma unroll
  for (int ro = 0; ro < R; ++ro) v[ro] = zero_src ? 0.0f : src[off[ro]];
  return ok_bits;
}
template <int R, bool FAST = false>
__device__ __forceinline__ void row_finish(const float (&v)[R], unsigned long long ok_bits, const int *oxo,
                                           const float *fr, float negM, float e_fill, int W, int delta, int rot_mode,
                                           int rot_len, bool use_mask, const UnaryMask &mk, int x, float (&o)[R]) {
  float e[R];
#pragma unroll
  for (int ro = 0; ro < R; ++ro) {
    float vv = v[ro];
    if (use_mask && masked(mk, ro - delta, x + oxo[ro], W)) vv = kLogZero;
    e[ro] = ((ok_bits >> ro) & 1ull) ? (FAST ? __expf(vv + negM) : expf(vv + negM)) : e_fill;
  }
  if (rot_mode == kRotFilter) {
    RotDispatch<R, (R % 2 == 1) ? R - 2 : R - 1>::run(rot_len, e, o, fr);
  } else if (rot_mode == kRotCopy) {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) o[ro] = e[ro];
  } else {
#pragma unroll
    for (int ro = 0; ro < R; ++ro) o[ro] = 0.0f;
  }
}

}  // namespace ps
