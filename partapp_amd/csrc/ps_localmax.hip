// findLocalMax (objectdetect_aux.cpp:193-251) on device grids.
//
// A point (d, y, x) of a [D][H][W] grid is kept when no spatial 8-neighbour
// in its slice is strictly greater and the slices d-1 and d+1 (where they
// exist) are strictly lower at (y, x).  With at most max_n points the
// reference returns them in its scan order (d, then x, then y); with more it
// keeps the max_n largest values, sorted by value (ties by scan order here;
// the reference's std::sort leaves their order unspecified).
//
//   k_lmax_mark    one thread per grid point: the test, then an atomic append
//                  of key = (order key of value) << 32 | (2^32-1 - scan index)
//   k_lmax_select  one workgroup per grid: at most max_n candidates -> sort
//                  by scan index; more -> radix-select the max_n largest keys
//                  (8 passes of 8 bits) and bitonic-sort them in LDS
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/partapp_ps.h"
#include "ps_plan.hpp"

namespace ps {

namespace {

struct LmGrid {
  long long off;       // grid start in the input array (elements)
  long long cand_off;  // start of its candidate buffer (keys)
  int D, H, W;
  int pad_;
};

constexpr int kSelThreads = 1024;
constexpr int kMaxLocalMax = 4096;

__device__ __forceinline__ unsigned okey(float f) {
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_inv(unsigned k) {
  unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__global__ __launch_bounds__(256) void k_lmax_clear(int *__restrict__ cnt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = 0;
}

__global__ __launch_bounds__(256) void k_lmax_mark(const float *__restrict__ grids, const LmGrid *__restrict__ gd,
                                                   unsigned long long *__restrict__ cand, int *__restrict__ count) {
  const LmGrid G = gd[blockIdx.y];
  const long long n = (long long)G.D * G.H * G.W;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int HW = G.H * G.W;
  const int d = (int)(e / HW);
  const int rem = (int)(e - (long long)d * HW);
  const int y = rem / G.W;
  const int x = rem - y * G.W;
  const float *g = grids + G.off;
  const float v = g[e];
  bool keep = true;
  for (int dy = -1; dy <= 1 && keep; ++dy)
    for (int dx = -1; dx <= 1 && keep; ++dx) {
      const int yy = y + dy, xx = x + dx;
      if ((dx | dy) != 0 && yy >= 0 && yy < G.H && xx >= 0 && xx < G.W && g[(long long)d * HW + yy * G.W + xx] > v)
        keep = false;
    }
  if (keep && d > 0) keep = g[e - HW] < v;
  if (keep && d < G.D - 1) keep = g[e + HW] < v;
  if (!keep) return;
  const unsigned scan = (unsigned)(((long long)d * G.W + x) * G.H + y);  // d, then x, then y
  const int slot = atomicAdd(count + blockIdx.y, 1);
  cand[G.cand_off + slot] = ((unsigned long long)okey(v) << 32) | (unsigned long long)(0xffffffffu - scan);
}

// bitonic sort of n (a power of two, <= 2 * kSelThreads... up to kMaxLocalMax)
// keys in LDS, descending
__device__ void bitonic_desc(unsigned long long *s, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int p = i ^ j;
        if (p > i) {
          const bool desc = (i & k) == 0;
          const unsigned long long a = s[i], b = s[p];
          if (desc ? (a < b) : (a > b)) {
            s[i] = b;
            s[p] = a;
          }
        }
      }
    }
  __syncthreads();
}

__global__ __launch_bounds__(kSelThreads) void k_lmax_select(const LmGrid *__restrict__ gd,
                                                             const unsigned long long *__restrict__ cand,
                                                             const int *__restrict__ count, int max_n,
                                                             ps_local_max *__restrict__ out, int *__restrict__ out_n) {
  __shared__ unsigned long long s[kMaxLocalMax];
  __shared__ unsigned hist[256];
  __shared__ unsigned long long prefix_s;
  __shared__ unsigned need_s;
  __shared__ int fill_s;
  const LmGrid G = gd[blockIdx.x];
  const unsigned long long *c = cand + G.cand_off;
  const int cnt = count[blockIdx.x];
  const int keep = cnt < max_n ? cnt : max_n;
  int n2 = 1;
  while (n2 < keep) n2 <<= 1;

  if (cnt <= max_n) {
    // all of them, in scan order: sort by the low word (2^32-1 - scan) descending
    for (int i = threadIdx.x; i < n2; i += blockDim.x) s[i] = i < cnt ? (c[i] << 32) | (c[i] >> 32) : 0ull;  // swap the halves
    __syncthreads();
    bitonic_desc(s, n2);
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) s[i] = (s[i] >> 32) | (s[i] << 32);  // back to (value, scan)
    __syncthreads();
  } else {
    // radix select of the max_n largest keys, most significant byte first
    if (threadIdx.x == 0) {
      prefix_s = 0ull;
      need_s = (unsigned)max_n;
    }
    unsigned long long mask = 0ull;
    for (int shift = 56; shift >= 0; shift -= 8) {
      for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
      __syncthreads();
      const unsigned long long prefix = prefix_s;
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const unsigned long long k = c[i];
        if ((k & mask) == prefix) atomicAdd(&hist[(unsigned)(k >> shift) & 255u], 1u);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned need = need_s, acc = 0u;
        int bin = 255;
        for (; bin > 0; --bin) {
          if (acc + hist[bin] >= need) break;
          acc += hist[bin];
        }
        need_s = need - acc;
        prefix_s = prefix | ((unsigned long long)bin << shift);
      }
      mask |= 0xffull << shift;
      __syncthreads();
    }
    // keys are distinct: exactly max_n of them are >= the selected key
    const unsigned long long thr = prefix_s;
    if (threadIdx.x == 0) fill_s = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const unsigned long long k = c[i];
      if (k >= thr) s[atomicAdd(&fill_s, 1)] = k;
    }
    __syncthreads();
    for (int i = keep + threadIdx.x; i < n2; i += blockDim.x) s[i] = 0ull;
    bitonic_desc(s, n2);
  }
  ps_local_max *o = out + (long long)blockIdx.x * max_n;
  for (int i = threadIdx.x; i < keep; i += blockDim.x) {
    const unsigned long long k = s[i];
    const unsigned scan = 0xffffffffu - (unsigned)(k & 0xffffffffull);
    ps_local_max r;
    const unsigned hw = (unsigned)G.H * (unsigned)G.W;
    r.d = (int)(scan / hw);
    const unsigned rem = scan - (unsigned)r.d * hw;
    r.x = (int)(rem / (unsigned)G.H);
    r.y = (int)(rem - (unsigned)r.x * (unsigned)G.H);
    r.score = okey_inv((unsigned)(k >> 32));
    o[i] = r;
  }
  if (threadIdx.x == 0) out_n[blockIdx.x] = keep;
}

}  // namespace

}  // namespace ps

using namespace ps;

extern "C" int ps_find_local_max(const float *grids_dev, const long long *offsets, const int *dims, int n_grids,
                                 int max_n, ps_local_max *out_dev, int *count_dev, void *stream) {
  if (!grids_dev || !offsets || !dims || n_grids < 1 || !out_dev || !count_dev)
    return set_error(PS_ERR_INVALID_ARG, "null argument or n_grids < 1");
  if (max_n < 1 || max_n > kMaxLocalMax) return set_error(PS_ERR_UNSUPPORTED, "max_n must be in [1, 4096]");
  std::vector<LmGrid> gd(n_grids);
  long long total = 0, max_n_el = 0;
  for (int g = 0; g < n_grids; ++g) {
    const int D = dims[3 * g], H = dims[3 * g + 1], W = dims[3 * g + 2];
    if (D < 1 || H < 1 || W < 1 || offsets[g] < 0) return set_error(PS_ERR_INVALID_ARG, "bad grid dims or offset");
    const long long n = (long long)D * H * W;
    if (n >= (1ll << 31)) return set_error(PS_ERR_UNSUPPORTED, "grid too large");
    gd[g].off = offsets[g];
    gd[g].cand_off = total;
    gd[g].D = D;
    gd[g].H = H;
    gd[g].W = W;
    gd[g].pad_ = 0;
    total += n;
    max_n_el = std::max(max_n_el, n);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // workspace: grid table, counters, one candidate slot per grid point
  const size_t o_cnt = (sizeof(LmGrid) * n_grids + 255) & ~(size_t)255;
  const size_t o_cand = (o_cnt + sizeof(int) * n_grids + 255) & ~(size_t)255;
  const size_t bytes = o_cand + sizeof(unsigned long long) * (size_t)total;
  void *ws = nullptr;
  if (hipMalloc(&ws, bytes) != hipSuccess) return set_error(PS_ERR_OUT_OF_MEMORY, "local max workspace");
  char *base = static_cast<char *>(ws);
  LmGrid *gdev = reinterpret_cast<LmGrid *>(base);
  int *cnt = reinterpret_cast<int *>(base + o_cnt);
  unsigned long long *cand = reinterpret_cast<unsigned long long *>(base + o_cand);
  hipError_t e = hipMemcpyAsync(gdev, gd.data(), sizeof(LmGrid) * n_grids, hipMemcpyHostToDevice, s);
  // (this entry point allocates its workspace, uploads the grid table from
  // pageable memory and synchronizes: it is synchronous and not capturable)
  if (e == hipSuccess) {  // counters cleared by a small kernel
    hipLaunchKernelGGL(k_lmax_clear, dim3((n_grids + 255) / 256), dim3(256), 0, s, cnt, n_grids);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    dim3 grid((unsigned)((max_n_el + 255) / 256), n_grids);
    hipLaunchKernelGGL(k_lmax_mark, grid, dim3(256), 0, s, grids_dev, gdev, cand, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lmax_select, dim3(n_grids), dim3(kSelThreads), 0, s, gdev, cand, cnt, max_n, out_dev,
                       count_dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(ws);
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("local max: ") + hipGetErrorString(e));
  return PS_OK;
}
