// Unary ingestion on the device: a part detector's score grids mapped to
// image coordinates, clipped and logged — the step right before the sweep.
//
//   PartApp::loadScoreGrid (partapp.cpp:866-899): each slice (scale,
//   rotation) of the score grid goes through transform_grid_fixed_size with
//   Tig = Ti2 T2g and NO_CLASS_VALUE = 0: a nearest scatter (TM_DIRECT, the
//   last source cell of the x-outer, y-inner scan wins, zero cells are not
//   written) or a bilinear gather (multi_array_transform.hpp:122-239);
//   then clip_scores_fill (objectdetect_aux.hpp:41-59: negative -> 1e-4) and
//   computeLogGrid (multi_array_op.hpp:154-166: 0 -> LOG_ZERO), as
//   findObjectImageRotJoints does (objectdetect_findrot.cpp:834-845).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/partapp_ps.h"
#include "ps_device.hpp"
#include "ps_plan.hpp"

namespace ps {
extern thread_local std::string g_last_error;

namespace {

constexpr int kIngThreads = 256;

// clip_scores_fill then computeLogGrid of one value
__device__ __forceinline__ float clip_log(float v) {
  if (v < 0.0f) v = 0.0001f;
  return v == 0.0f ? kLogZero : logf(v);
}

}  // namespace

__global__ __launch_bounds__(kIngThreads) void k_ingest_clear(unsigned long long *keys, long long n) {
  for (long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kIngThreads)
    keys[i] = 0ull;
}

// TM_DIRECT: source cell (x1, y1) of slice s lands on the nearest output cell
// of Tig (x1, y1); the key (scan position + 1, value bits) keeps the last.
__global__ __launch_bounds__(kIngThreads) void k_ingest_scatter(const float *__restrict__ grid,
                                                                const double *__restrict__ mats, int n, int gh,
                                                                int gw, int H, int W,
                                                                unsigned long long *__restrict__ keys) {
  const long long cells = (long long)gh * gw;
  const long long total = n * cells;
  for (long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x; i < total;
       i += (long long)gridDim.x * kIngThreads) {
    const float v = grid[i];
    if (v == 0.0f) continue;  // NO_CLASS_VALUE is not written
    const int s = (int)(i / cells);
    const int rem = (int)(i - s * cells);
    const int y1 = rem / gw;
    const int x1 = rem - y1 * gw;
    const double *m = mats + 6 * s;
    const double x3 = m[0] * (double)x1 + m[1] * (double)y1 + m[2];
    const double y3 = m[3] * (double)x1 + m[4] * (double)y1 + m[5];
    const int ix = (int)floor(x3 + 0.5);
    const int iy = (int)floor(y3 + 0.5);
    if (ix >= 0 && ix < W && iy >= 0 && iy < H) {
      const unsigned scan = (unsigned)x1 * (unsigned)gh + (unsigned)y1 + 1u;
      const unsigned long long key = ((unsigned long long)scan << 32) | (unsigned long long)__float_as_uint(v);
      atomicMax(keys + ((long long)s * H + iy) * W + ix, key);
    }
  }
}

__global__ __launch_bounds__(kIngThreads) void k_ingest_finish(const unsigned long long *__restrict__ keys,
                                                               long long n, float *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kIngThreads) {
    const unsigned long long k = keys[i];
    out[i] = clip_log(k ? __uint_as_float((unsigned)(k & 0xffffffffull)) : 0.0f);
  }
}

// TM_BILINEAR: output pixel (x3, y3) samples its slice at inverse(Tig) (x3, y3)
// (within 10 float epsilons of a grid point: that point; last row / column
// and outside: NO_CLASS_VALUE)
__global__ __launch_bounds__(kIngThreads) void k_ingest_gather(const float *__restrict__ grid,
                                                               const double *__restrict__ mats, int n, int gh,
                                                               int gw, int H, int W, float *__restrict__ out) {
  const long long cells = (long long)H * W;
  const long long total = n * cells;
  const float eps10 = 10.0f * FLT_EPSILON;
  for (long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x; i < total;
       i += (long long)gridDim.x * kIngThreads) {
    const int s = (int)(i / cells);
    const int rem = (int)(i - s * cells);
    const int y3 = rem / W;
    const int x3 = rem - y3 * W;
    const double *m = mats + 6 * s;
    const double x1 = m[0] * (double)x3 + m[1] * (double)y3 + m[2];
    const double y1 = m[3] * (double)x3 + m[4] * (double)y3 + m[5];
    const int ix = (int)floor(x1);
    const int iy = (int)floor(y1);
    float r = 0.0f;
    if (ix >= 0 && ix < gw && iy >= 0 && iy < gh) {
      const float a = (float)(x1 - (double)ix);
      const float b = (float)(y1 - (double)iy);
      const float *p = grid + (long long)s * gh * gw + (long long)iy * gw + ix;
      if (a < eps10 && b < eps10) {
        r = p[0];
      } else if (ix < gw - 1 && iy < gh - 1) {
        const float t0 = (1.0f - b) * (1.0f - a) * p[0];
        const float t1 = (1.0f - b) * a * p[1];
        const float t2 = b * (1.0f - a) * p[gw];
        const float t3 = b * a * p[gw + 1];
        r = t0 + t1 + t2 + t3;
      }
    }
    out[i] = clip_log(r);
  }
}

}  // namespace ps

using namespace ps;

extern "C" int ps_ingest_scores(const float *score_grid_dev, int n, int grid_height, int grid_width,
                                const double *Tig, int interpolate, int height, int width, float *log_unary_dev,
                                void *stream) {
  if (!score_grid_dev || !Tig || !log_unary_dev) return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (n < 1 || grid_height < 1 || grid_width < 1 || height < 1 || width < 1)
    return set_error(PS_ERR_INVALID_ARG, "counts and sizes must be >= 1");
  if ((long long)grid_height * grid_width >= (1ll << 31) || (long long)height * width >= (1ll << 31) ||
      (long long)grid_width * grid_height + 1 >= (1ll << 32))
    return set_error(PS_ERR_UNSUPPORTED, "grids must have fewer than 2^31 cells per slice");
  for (int s = 0; s < n; ++s) {
    const double *T = Tig + 9 * s;
    if (interpolate && !(T[0] * T[4] - T[3] * T[1] != 0.0))
      return set_error(PS_ERR_INVALID_ARG, "Tig of slice " + std::to_string(s) + " is singular");
  }
  int devcount = 0;
  if (hipGetDeviceCount(&devcount) != hipSuccess || devcount == 0)
    return set_error(PS_ERR_NO_DEVICE, "no HIP device");
  std::vector<double> mats(6 * (size_t)n);
  for (int s = 0; s < n; ++s) warp_matrix(Tig + 9 * s, interpolate != 0, mats.data() + 6 * s);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long out_n = (long long)n * height * width;
  const size_t mat_bytes = (mats.size() * sizeof(double) + 255) & ~(size_t)255;
  const size_t key_bytes = interpolate ? 0 : (size_t)out_n * sizeof(unsigned long long);
  void *ws = nullptr;
  hipError_t e = hipMalloc(&ws, mat_bytes + key_bytes);
  if (e != hipSuccess) return set_error(PS_ERR_OUT_OF_MEMORY, "ingest workspace");
  double *mats_dev = static_cast<double *>(ws);
  unsigned long long *keys = reinterpret_cast<unsigned long long *>(static_cast<char *>(ws) + mat_bytes);
  e = hipMemcpyAsync(mats_dev, mats.data(), mats.size() * sizeof(double), hipMemcpyHostToDevice, st);
  auto blocks = [](long long m) {
    long long g = (m + kIngThreads - 1) / kIngThreads;
    return (unsigned)(g < 8192 ? (g < 1 ? 1 : g) : 8192);
  };
  if (e == hipSuccess) {
    if (interpolate) {
      hipLaunchKernelGGL(k_ingest_gather, dim3(blocks(out_n)), dim3(kIngThreads), 0, st, score_grid_dev, mats_dev, n,
                         grid_height, grid_width, height, width, log_unary_dev);
    } else {
      const long long in_n = (long long)n * grid_height * grid_width;
      hipLaunchKernelGGL(k_ingest_clear, dim3(blocks(out_n)), dim3(kIngThreads), 0, st, keys, out_n);
      hipLaunchKernelGGL(k_ingest_scatter, dim3(blocks(in_n)), dim3(kIngThreads), 0, st, score_grid_dev, mats_dev, n,
                         grid_height, grid_width, height, width, keys);
      hipLaunchKernelGGL(k_ingest_finish, dim3(blocks(out_n)), dim3(kIngThreads), 0, st, keys, out_n,
                         log_unary_dev);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(ws);
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("ingest: ") + hipGetErrorString(e));
  return PS_OK;
}

// ------------------------------------------------------------- extra unaries
// The additions findObjectImageRotJoints makes to the log unaries before the
// sweep (objectdetect_findrot.cpp:896-948), in place on device grids, with
// the reference's float / double mix (objectdetect_icps.cpp is compiled with
// `using namespace std`: log(float) is the float log, exp and pow(float, int)
// work in double, and `float += double` adds in double and rounds once).

namespace ps {
namespace {

// setTorsoPosPrior (objectdetect_icps.cpp:137-191): one thread per pixel
// forms the clipped log prior once, then adds it to every slice.
__global__ __launch_bounds__(kIngThreads) void k_torso_prior(float *g, int n_slices, int H, int W, float mu_x,
                                                             float mu_y, float var_x, float var_y, float weight) {
  const long long hw = (long long)H * W;
  const long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x;
  if (i >= hw) return;
  const int iy = (int)(i / W), ix = (int)(i - (long long)iy * W);
  const float img_c_x = (float)(0.5 * W), img_c_y = (float)(0.5 * H);
  const float ix_rel = img_c_x - (float)ix;
  const float iy_rel = img_c_y - (float)iy;
  // pow(d, 2) of a float d is exact in double
  const double dx = (double)(ix_rel - mu_x), dy = (double)(iy_rel - mu_y);
  const float score_x = (float)exp(-0.5 * (dx * dx) / (double)var_x);
  const float score_y = (float)exp(-0.5 * (dy * dy) / (double)var_y);
  float score = weight * score_x * score_y;
  if ((double)score < 1e-4) score = 1e-4f;
  const double pred = (double)logf(score);
  for (int k = 0; k < n_slices; ++k) {
    float *p = g + (long long)k * hw + i;
    *p = (float)((double)*p + pred);
  }
}

// addLoadDPMScore's accumulation (:469-484, raw scores), addDPMScore
// (:490-524, log grids) and the torso path's computeLogGrid + addDPMScore
// (objectdetect_findrot.cpp:896-903): one thread per pixel, all slices
__global__ __launch_bounds__(kIngThreads) void k_dpm_add(float *g, int S, int R, int H, int W, const float *dpm,
                                                         int n_dpm, float weight, int mode, double log_clip) {
  const long long hw = (long long)H * W;
  const long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x;
  if (i >= hw) return;
  for (int s = 0; s < S; ++s)
    for (int r = 0; r < R; ++r) {
      float *p = g + ((long long)s * R + r) * hw + i;
      if (mode == PS_DPM_RAW) {
        const float val = dpm[(long long)(n_dpm == R ? r : 0) * hw + i];
        const double add = (double)val > 1e-4 ? (double)(weight * logf(val)) : log_clip;
        *p = (float)((double)*p + add);
      } else {
        float lv = dpm[(long long)(n_dpm > 1 ? r : 0) * hw + i];
        if (mode == PS_DPM_LOG_OF_RAW) lv = lv == 0.0f ? kLogZero : logf(lv);
        *p = *p + weight * lv;
      }
    }
}

// addExtraUnary (:526-548)
__global__ __launch_bounds__(kIngThreads) void k_add_scaled(float *g, const float *e, long long n, float weight) {
  for (long long i = (long long)blockIdx.x * kIngThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kIngThreads)
    g[i] = g[i] + weight * e[i];
}

unsigned pixel_blocks(long long hw) { return (unsigned)((hw + kIngThreads - 1) / kIngThreads); }

}  // namespace
}  // namespace ps

extern "C" int ps_add_torso_pos_prior(float *root_unary_dev, int n_slices, int height, int width,
                                      const double params[4], float weight, void *stream) {
  if (!root_unary_dev || !params) return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (n_slices < 1 || height < 1 || width < 1) return set_error(PS_ERR_INVALID_ARG, "counts and sizes must be >= 1");
  if ((long long)height * width >= (1ll << 40)) return set_error(PS_ERR_UNSUPPORTED, "grid too large");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_torso_prior, dim3(pixel_blocks((long long)height * width)), dim3(kIngThreads), 0, st,
                     root_unary_dev, n_slices, height, width, (float)params[0], (float)params[1], (float)params[2],
                     (float)params[3], weight);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("torso prior: ") + hipGetErrorString(e));
  return PS_OK;
}

extern "C" int ps_add_dpm_score(float *part_unary_dev, int S, int R, int height, int width, const float *dpm_dev,
                                int n_dpm, float weight, int mode, void *stream) {
  if (!part_unary_dev || !dpm_dev) return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (S < 1 || R < 1 || height < 1 || width < 1 || n_dpm < 1)
    return set_error(PS_ERR_INVALID_ARG, "counts and sizes must be >= 1");
  if (mode != PS_DPM_RAW && mode != PS_DPM_LOG && mode != PS_DPM_LOG_OF_RAW)
    return set_error(PS_ERR_INVALID_ARG, "unknown DPM mode " + std::to_string(mode));
  if (mode != PS_DPM_RAW) {
    if (!(weight > 0)) return set_error(PS_ERR_INVALID_ARG, "dpm_weight must be > 0 (objectdetect_icps.cpp:501)");
    if (n_dpm > 1 && n_dpm < R) return set_error(PS_ERR_INVALID_ARG, "a rotation-dependent DPM prior needs R grids");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_dpm_add, dim3(pixel_blocks((long long)height * width)), dim3(kIngThreads), 0, st,
                     part_unary_dev, S, R, height, width, dpm_dev, n_dpm, weight, mode, std::log(1e-4));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("dpm score: ") + hipGetErrorString(e));
  return PS_OK;
}

extern "C" int ps_add_extra_unary(float *log_unary_dev, const float *extra_dev, long long n, float weight,
                                  void *stream) {
  if (!log_unary_dev || !extra_dev) return set_error(PS_ERR_INVALID_ARG, "null argument");
  if (n < 0) return set_error(PS_ERR_INVALID_ARG, "n must be >= 0");
  if (n == 0) return PS_OK;
  long long g = (n + kIngThreads - 1) / kIngThreads;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_add_scaled, dim3((unsigned)(g < 65536 ? g : 65536)), dim3(kIngThreads), 0, st, log_unary_dev,
                     extra_dev, n, weight);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("extra unary: ") + hipGetErrorString(e));
  return PS_OK;
}
