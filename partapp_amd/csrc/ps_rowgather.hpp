// The circular rotation filter of the FP32 row pass (objectdetect_findrot.cpp:
// 377-420) on a register column, one instantiation per tap count.
#pragma once

#include <hip/hip_runtime.h>

#include "ps_device.hpp"
#include "ps_internal.hpp"

namespace ps {

// Circular rotation filter with LR taps on a register column
// (objectdetect_findrot.cpp:377-420, multi_array_filter.hpp:115-155):
// out[r] = sum_j in[(r - n + j) mod R] f[j], taps in ascending order.  Outputs
// LO .. HI - 1 are stored as they are formed, output r at out[(r - LO) pitch].
template <int R, int LR, int LO, int HI>
__device__ __forceinline__ void rot_filter_fixed(const float (&e)[R], const float *fr, float *out, int pitch) {
  constexpr int n = (LR - 1) / 2;
  float f[LR];
#pragma unroll
  for (int j = 0; j < LR; ++j) f[j] = fr[j];
#pragma unroll
  for (int ro = LO; ro < HI; ++ro) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < LR; ++j) acc = fmaf(e[((ro - n + j) % R + R) % R], f[j], acc);
    out[(ro - LO) * pitch] = acc;
  }
}
template <int R, int LR, int LO, int HI>
struct RotDispatch {
  static __device__ __forceinline__ void run(int lr, const float (&e)[R], const float *fr, float *out, int pitch) {
    if (lr == LR) rot_filter_fixed<R, LR, LO, HI>(e, fr, out, pitch);
    else RotDispatch<R, LR - 2, LO, HI>::run(lr, e, fr, out, pitch);
  }
};
template <int R, int LO, int HI>
struct RotDispatch<R, -1, LO, HI> {
  static __device__ __forceinline__ void run(int, const float (&)[R], const float *, float *out, int pitch) {
#pragma unroll
    for (int r = 0; r < HI - LO; ++r) out[r * pitch] = 0.0f;
  }
};

}  // namespace ps
