// Eigenbasis path of a message whose covariance has off-diagonal terms
// (gaussFilter2dOffset with offset 0, multi_array_filter.hpp:334-369, called
// per rotation slice by computeRotJointMarginal, objectdetect_findrot.cpp:
// 422-429).  The row pass has already written the rotation-filtered grid
// T[R][H][W] (exp(b - M) after the forward shift and the circular filter).
// Per slice:
//   k_od_scatter / k_od_gather  rotate T into the eigenbasis grid A[th][tw]:
//       nearest scatter (TM_DIRECT, sparse models: the last source cell of the
//       x-outer, y-inner scan wins, zero cells are not written) or a bilinear
//       gather (TM_BILINEAR) (multi_array_transform.hpp:122-239)
//   k_od_xfilter, k_od_yfilter  the truncated separable filter with the
//       eigenvalue taps (multi_array_filter.hpp:211-321)
//   k_od_back  bilinear sample back to H x W, log, +M, the inverse shift and
//       the caller's accumulation (the column pass's epilogue)
// Coordinates are mapped in double with the reference's operation order; the
// bilinear weights are float.  Every launch covers the off-diagonal messages
// of one level; a job's threads walk its cells with a grid-stride loop.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ps_device.hpp"
#include "ps_internal.hpp"

namespace ps {

namespace {

constexpr int kOdThreads = 256;

__device__ __forceinline__ unsigned long long *od_keys(const Bases &b, const MsgJob &J) {
  return reinterpret_cast<unsigned long long *>(mptr(b, J.od_buf));
}
__device__ __forceinline__ float *od_a(const Bases &b, const MsgJob &J) { return mptr(b, J.od_buf); }
__device__ __forceinline__ float *od_x(const Bases &b, const MsgJob &J, int R) {
  return mptr(b, J.od_buf) + 2ll * R * J.th * J.tw;
}

// TM_BILINEAR sample of a [h][w] grid (row pitch `pitch`) at (x1, y1)
// (multi_array_transform.hpp:218-236): within 10 float epsilons of a grid
// point the point itself, the last row and column (no right / lower
// neighbour) and outside: 0
__device__ __forceinline__ float bilinear_at(const float *g, int h, int w, int pitch, double x1, double y1) {
  const int ix = (int)floor(x1);
  const int iy = (int)floor(y1);
  float r = 0.0f;
  if (ix >= 0 && ix < w && iy >= 0 && iy < h) {
    const float eps10 = 10.0f * FLT_EPSILON;
    const float a = (float)(x1 - (double)ix);
    const float bb = (float)(y1 - (double)iy);
    const float *p = g + (long long)iy * pitch + ix;
    if (a < eps10 && bb < eps10) {
      r = p[0];
    } else if (ix < w - 1 && iy < h - 1) {
      const float t0 = (1.0f - bb) * (1.0f - a) * p[0];
      const float t1 = (1.0f - bb) * a * p[1];
      const float t2 = bb * (1.0f - a) * p[pitch];
      const float t3 = bb * a * p[pitch + 1];
      r = t0 + t1 + t2 + t3;
    }
  }
  return r;
}

}  // namespace

// Zero the scatter keys of the direct jobs.
__global__ __launch_bounds__(kOdThreads) void k_od_clear(Bases b, const MsgJob *__restrict__ jobs, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  if (!J.od_direct) return;
  const long long n = (long long)R * J.th * J.tw;
  unsigned long long *k = od_keys(b, J);
  for (long long i = (long long)blockIdx.x * kOdThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kOdThreads)
    k[i] = 0ull;
}

// TM_DIRECT: source cell (x1, y1) of slice r goes to the nearest cell of
// T31 (x1, y1).  The key (scan position + 1, value bits) keeps the cell that
// comes last in the reference's x-outer, y-inner scan.
__global__ __launch_bounds__(kOdThreads) void k_od_scatter(Bases b, const MsgJob *__restrict__ jobs, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  if (!J.od_direct) return;
  const int H = J.H, W = J.W, th = J.th, tw = J.tw;
  const long long HW = (long long)H * W;
  const long long n = R * HW;
  const float *T = cptr(b, J.tmp);
  unsigned long long *K = od_keys(b, J);
  const double m00 = J.t_fwd[0], m01 = J.t_fwd[1], m02 = J.t_fwd[2];
  const double m10 = J.t_fwd[3], m11 = J.t_fwd[4], m12 = J.t_fwd[5];
  const int tp = J.tpitch;
  for (long long i = (long long)blockIdx.x * kOdThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kOdThreads) {
    const int r = (int)(i / HW);
    const int rem = (int)(i - r * HW);
    const int y1 = rem / W;
    const int x1 = rem - y1 * W;
    const float v = T[((long long)r * H + y1) * tp + x1];
    if (v == 0.0f) continue;  // the default value is not written
    const double x3 = m00 * (double)x1 + m01 * (double)y1 + m02;
    const double y3 = m10 * (double)x1 + m11 * (double)y1 + m12;
    const int ix = (int)floor(x3 + 0.5);
    const int iy = (int)floor(y3 + 0.5);
    if (ix >= 0 && ix < tw && iy >= 0 && iy < th) {
      const unsigned scan = (unsigned)x1 * (unsigned)H + (unsigned)y1 + 1u;
      const unsigned long long key = ((unsigned long long)scan << 32) | (unsigned long long)__float_as_uint(v);
      atomicMax(K + ((long long)r * th + iy) * tw + ix, key);
    }
  }
}

// TM_BILINEAR: eigenbasis cell (x3, y3) samples slice r at T13 (x3, y3).
__global__ __launch_bounds__(kOdThreads) void k_od_gather(Bases b, const MsgJob *__restrict__ jobs, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  if (J.od_direct) return;
  const int H = J.H, W = J.W, th = J.th, tw = J.tw;
  const long long cells = (long long)th * tw;
  const long long n = R * cells;
  const float *T = cptr(b, J.tmp);
  float *A = od_a(b, J);
  const double m00 = J.t_fwd[0], m01 = J.t_fwd[1], m02 = J.t_fwd[2];
  const double m10 = J.t_fwd[3], m11 = J.t_fwd[4], m12 = J.t_fwd[5];
  for (long long i = (long long)blockIdx.x * kOdThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kOdThreads) {
    const int r = (int)(i / cells);
    const int rem = (int)(i - r * cells);
    const int y3 = rem / tw;
    const int x3 = rem - y3 * tw;
    const double x1 = m00 * (double)x3 + m01 * (double)y3 + m02;
    const double y1 = m10 * (double)x3 + m11 * (double)y3 + m12;
    A[i] = bilinear_at(T + (long long)r * H * J.tpitch, H, W, J.tpitch, x1, y1);
  }
}

// x filter of the eigenbasis grid: X = A * fx, zero padded, taps ascending
// (a zero-padded tap leaves the running sum unchanged, so it is skipped).
__global__ __launch_bounds__(kOdThreads) void k_od_xfilter(Bases b, const MsgJob *__restrict__ jobs,
                                                           const float *__restrict__ taps, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  const int th = J.th, tw = J.tw;
  const long long n = (long long)R * th * tw;
  const int k = J.kxe, L = 2 * J.kxe + 1;
  const float *fx = taps + J.taps_xe;
  const bool direct = J.od_direct != 0;
  const float *A = od_a(b, J);
  const unsigned long long *K = od_keys(b, J);
  float *X = od_x(b, J, R);
  for (long long i = (long long)blockIdx.x * kOdThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kOdThreads) {
    const long long row = i / tw;
    const int x = (int)(i - row * tw);
    const long long r0 = row * tw;
    float acc = 0.0f;
    for (int t = 0; t < L; ++t) {
      const int u = x - k + t;
      if (u < 0 || u >= tw) continue;
      float v;
      if (direct) {
        const unsigned long long key = K[r0 + u];
        v = key ? __uint_as_float((unsigned)(key & 0xffffffffull)) : 0.0f;
      } else {
        v = A[r0 + u];
      }
      acc = fmaf(v, fx[t], acc);
    }
    X[i] = acc;
  }
}

// y filter of the eigenbasis grid: A = X * fy.
__global__ __launch_bounds__(kOdThreads) void k_od_yfilter(Bases b, const MsgJob *__restrict__ jobs,
                                                           const float *__restrict__ taps, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  const int th = J.th, tw = J.tw;
  const long long cells = (long long)th * tw;
  const long long n = R * cells;
  const int k = J.kye, L = 2 * J.kye + 1;
  const float *fy = taps + J.taps_ye;
  const float *X = od_x(b, J, R);
  float *A = od_a(b, J);
  for (long long i = (long long)blockIdx.x * kOdThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kOdThreads) {
    const int r = (int)(i / cells);
    const int rem = (int)(i - r * cells);
    const int y = rem / tw;
    const int x = rem - y * tw;
    const float *col = X + (long long)r * cells + x;
    float acc = 0.0f;
    for (int t = 0; t < L; ++t) {
      const int yy = y - k + t;
      if (yy < 0 || yy >= th) continue;
      acc = fmaf(col[(long long)yy * tw], fy[t], acc);
    }
    A[i] = acc;
  }
}

// Back to the grid: pixel (x, y) of slice r reads the smoothed slice at
// (x + inv_ox[r], y + inv_oy[r]) (the inverse nearest warp, LOG_ZERO outside,
// with a split warp's step past its break,
// objectdetect_findrot.cpp:437-448), which is the bilinear sample of the
// eigenbasis grid at T_back of that pixel, log, +M (:432-435); then the
// caller's accumulation.
__global__ __launch_bounds__(kOdThreads) void k_od_back(Bases b, const MsgJob *__restrict__ jobs,
                                                        const int *__restrict__ shifts, int R) {
  const MsgJob &J = jobs[blockIdx.y];
  if (skip_inst(b, J.inst)) return;
  const int H = J.H, W = J.W, th = J.th, tw = J.tw;
  const long long HW = (long long)H * W;
  const long long n = R * HW;
  const long long i0 = (long long)blockIdx.x * kOdThreads;
  if (i0 >= n) return;  // uniform over the block
  const long long i = i0 + threadIdx.x;
  const bool ok = i < n;
  const int r = ok ? (int)(i / HW) : 0;
  const int rem = ok ? (int)(i - r * HW) : 0;
  const int y = rem / W;
  const int x = rem - y * W;
  const int *sh = shifts + J.shifts;
  int xs = x + sh[2 * R + r];
  int ys = y + sh[3 * R + r];
  if (J.split) {  // the inverse offsets change once along x / y: breaks and steps from 12 R on
    if (x >= sh[12 * R + r]) xs += sh[13 * R + r];
    if (y >= sh[14 * R + r]) ys += sh[15 * R + r];
  }
  const float M = key2f(b.max_keys[J.src_key]);
  float *dst = mptr(b, J.dst);
  const float *dun = J.dst_unary.buf >= 0 ? cptr(b, J.dst_unary) : nullptr;
  float *dst2 = J.dst2.buf >= 0 ? mptr(b, J.dst2) : nullptr;
  const float *bel = J.dst_bel.buf >= 0 ? cptr(b, J.dst_bel) : nullptr;
  // element (r, y, x) of a grid at its row pitch (i is the dense one)
  const long long ry = (long long)r * H + y;
  auto at = [&](const GridRef &g) { return ry * g.pitch + x; };
  DstAcc da;
  if (ok) {
    float msg = kLogZero;
    if (xs >= 0 && xs < W && ys >= 0 && ys < H) {
      const double xe = J.t_back[0] * (double)xs + J.t_back[1] * (double)ys + J.t_back[2];
      const double ye = J.t_back[3] * (double)xs + J.t_back[4] * (double)ys + J.t_back[5];
      const float v = bilinear_at(od_a(b, J) + (long long)r * th * tw, th, tw, tw, xe, ye);
      msg = log0(v) + M;
    }
    const bool need_u = (J.dst_kind == kDstUpNode && J.dst_add_unary) || J.dst_kind == kDstDown;
    const bool need_b = J.dst_kind == kDstDown && J.dst_base == 0;
    const float pu = need_u ? dun[at(J.dst_unary)] : 0.0f;
    const float pb = need_b ? bel[at(J.dst_bel)] : 0.0f;
    const bool xmask = ((J.dst_mask.rot_bits >> r) & 1ull) || x < J.dst_mask.strip || x >= W - J.dst_mask.strip;
    dst_store(J, dst, dst2, at(J.dst), dst2 ? at(J.dst2) : 0, (unsigned)i, xmask, msg, pu, pb, da);
  }
  dst_flush(b, J, dst2 != nullptr, da, threadIdx.x & 63);
}

// ------------------------------------------------------------ host launcher

// The eigenbasis stages of n_jobs off-diagonal messages; max_cells bounds
// R * th * tw and max_hw H * W over them.
hipError_t launch_offdiag(const Bases &b, const MsgJob *jobs, int n_jobs, const float *taps, const int *shifts, int R,
                          long long max_cells, long long max_hw, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  const dim3 blk(kOdThreads);
  auto blocks = [](long long n) {
    long long g = (n + kOdThreads - 1) / kOdThreads;
    return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
  };
  const unsigned gc = blocks(max_cells);
  const unsigned gs = blocks((long long)R * max_hw);
  hipLaunchKernelGGL(k_od_clear, dim3(gc, n_jobs), blk, 0, s, b, jobs, R);
  hipLaunchKernelGGL(k_od_scatter, dim3(gs, n_jobs), blk, 0, s, b, jobs, R);
  hipLaunchKernelGGL(k_od_gather, dim3(gc, n_jobs), blk, 0, s, b, jobs, R);
  hipLaunchKernelGGL(k_od_xfilter, dim3(gc, n_jobs), blk, 0, s, b, jobs, taps, R);
  hipLaunchKernelGGL(k_od_yfilter, dim3(gc, n_jobs), blk, 0, s, b, jobs, taps, R);
  // one thread per output element: the epilogue's wave reductions need every
  // lane of a wave, so this launch is not grid-stride
  const long long nb = (R * max_hw + kOdThreads - 1) / kOdThreads;
  hipLaunchKernelGGL(k_od_back, dim3((unsigned)nb, n_jobs), blk, 0, s, b, jobs, shifts, R);
  return hipGetLastError();
}

}  // namespace ps
