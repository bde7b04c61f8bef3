// Device-side generator of the synthetic log unaries, the same recipe as
// ps_synth_unary (ps_synth.cpp): counter-based SplitMix64 streams, uniforms
// in (0, 1], Box-Muller in double, N(-3, 0.5) plus three amplitude-4,
// std-6 px bumps per part in their rotation bin.  For batches too large to
// generate on the host (hundreds of instances, tens of GB); the values match
// the host generator up to the last bit of the device's double-precision
// log / sqrt / sin / cos.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/partapp_ps.h"
#include "ps_plan.hpp"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kTagBumps = 0x62756d7073ull;  // "bumps"
constexpr uint64_t kTagUnary = 0x756e617279ull;  // "unary"
constexpr int kBumps = 3;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_at(uint64_t state, uint64_t i) { return mix64(state + (i + 1) * kGamma); }
__host__ __device__ inline double u01(uint64_t z) { return (double)((z >> 11) + 1) * 0x1.0p-53; }
__host__ __device__ inline uint64_t key(uint64_t seed, uint64_t tag, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed ^ tag) + a) + b);
}

// bump positions per part: x, y, rotation (host, as ps_synth_unary)
void bumps(uint64_t seed, int P, int R, int H, int W, std::vector<int> &b) {
  b.resize((size_t)P * kBumps * 3);
  for (int p = 0; p < P; ++p) {
    const uint64_t state = key(seed, kTagBumps, (uint64_t)p, 0);
    uint64_t i = 0;
    auto index = [&](int n) {
      int v = (int)std::floor(u01(stream_at(state, i++)) * n);
      return v >= n ? n - 1 : v;
    };
    for (int k = 0; k < kBumps; ++k) {
      const int q = (p * kBumps + k) * 3;
      b[q] = index(W);
      b[q + 1] = index(H);
      b[q + 2] = index(R);
    }
  }
}

__global__ __launch_bounds__(256) void k_synth_unary(uint64_t seed, int P, int S, int R, int H, int W,
                                                     const int *__restrict__ bump, float *__restrict__ out) {
  const long long plane = (long long)H * W, n = (long long)R * plane;
  const long long total = (long long)P * S * n;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const long long ps = i / n;
    const long long e = i - ps * n;  // index inside the [R][H][W] grid
    const int p = (int)(ps / S), s = (int)(ps - (long long)p * S);
    const int r = (int)(e / plane);
    const long long rem = e - (long long)r * plane;
    const int y = (int)(rem / W), x = (int)(rem - (long long)y * W);
    const uint64_t state = key(seed, kTagUnary, (uint64_t)p, (uint64_t)s);
    const long long pair = e >> 1;
    const double u1 = u01(stream_at(state, 2 * pair));
    const double u2 = u01(stream_at(state, 2 * pair + 1));
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 2.0 * M_PI * u2;
    const double z = (e & 1) ? rad * sin(th) : rad * cos(th);
    double v = -3.0 + 0.5 * z;
    for (int k = 0; k < kBumps; ++k) {
      const int *bq = bump + (p * kBumps + k) * 3;
      if (bq[2] != r) continue;
      const double dx = x - bq[0], dy = y - bq[1];
      v += 4.0 * exp(-(dx * dx + dy * dy) / (2.0 * 36.0));
    }
    out[i] = (float)v;
  }
}

}  // namespace

extern "C" int ps_synth_unary_dev(unsigned long long seed, int P, int S, int R, int H, int W, float *unary_dev,
                                  void *stream) {
  if (P < 1 || S < 1 || R < 1 || H < 1 || W < 1 || !unary_dev)
    return ps::set_error(PS_ERR_INVALID_ARG, "bad argument");
  std::vector<int> b;
  bumps(seed, P, R, H, W, b);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int *bd = nullptr;
  hipError_t e = hipMalloc(&bd, b.size() * sizeof(int));
  if (e != hipSuccess) return ps::set_error(PS_ERR_OUT_OF_MEMORY, "bump table");
  e = hipMemcpyAsync(bd, b.data(), b.size() * sizeof(int), hipMemcpyHostToDevice, s);
  const long long total = (long long)P * S * R * H * W;
  long long g = (total + 255) / 256;
  if (g > 16384) g = 16384;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_synth_unary, dim3((unsigned)g), dim3(256), 0, s, (uint64_t)seed, P, S, R, H, W, bd,
                       unary_dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(bd);
  if (e != hipSuccess) return ps::set_error(PS_ERR_HIP, std::string("synth: ") + hipGetErrorString(e));
  return PS_OK;
}
