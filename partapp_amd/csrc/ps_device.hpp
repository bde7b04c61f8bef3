// Device helpers shared by the sweep kernels (ps_kernels.hip) and the
// eigenbasis path of off-diagonal covariances (ps_offdiag.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ps_internal.hpp"

namespace ps {

// order-preserving 32-bit key of a float (larger float, larger key)
__device__ __forceinline__ unsigned f2key(float f) {
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned k) {
  unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}
// argmax key: larger score wins, then the smaller flat index (the first
// maximum of a C-order scan, objectdetect_findrot.cpp:261-266)
__device__ __forceinline__ unsigned long long amax_key(float v, unsigned idx) {
  return ((unsigned long long)f2key(v) << 32) | (unsigned long long)(0xffffffffu - idx);
}

// The block of a 2-D grid (x fastest) renumbered so that each of the 8 XCDs,
// to which the dispatcher hands workgroups round-robin in launch order, gets
// one contiguous range of logical blocks.  Neighbouring tiles (which share the
// cache lines at their edges when grid rows are not 128-byte aligned) then run
// at the same time behind the same L2.
__device__ __forceinline__ void xcd_block(unsigned &bx, unsigned &by) {
  const unsigned n = gridDim.x * gridDim.y;
  const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned xcd = b & 7u, k = b >> 3;
  const unsigned q = n >> 3, rem = n & 7u;
  const unsigned l = xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
  by = l / gridDim.x;
  bx = l - by * gridDim.x;
}

__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// An instance left out of this execution (Bases::active names the instances
// that run): every block of its jobs exits at once.
__device__ __forceinline__ bool skip_inst(const Bases &b, int inst) { return b.active && b.active[inst] == 0u; }

// Top-2 value keys of a wave's outputs (per lane k1 >= k2, every state seen
// by one lane only), merged across the wave and then into a slot's
// (best << 32) | second word with a compare-and-swap loop; the word only
// grows, so the loop ends.  Every lane of the wave calls it.
__device__ __forceinline__ void top2_flush(unsigned long long *slot, unsigned k1, unsigned k2, int lane) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned o1 = (unsigned)__shfl_xor((int)k1, o, 64), o2 = (unsigned)__shfl_xor((int)k2, o, 64);
    k2 = max(max(k2, o2), min(k1, o1));
    k1 = max(k1, o1);
  }
  if (lane == 0 && k1) {
    unsigned long long old = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
      const unsigned a1 = (unsigned)(old >> 32), a2 = (unsigned)old;
      const unsigned long long nw = ((unsigned long long)max(a1, k1) << 32) | max(max(a2, k2), min(a1, k1));
      if (nw == old) break;
      const unsigned long long prev = atomicCAS(slot, old, nw);
      if (prev == old) break;
      old = prev;
    }
  }
}
// one more value key into a lane's top-2
__device__ __forceinline__ void top2_add(unsigned &k1, unsigned &k2, unsigned k) {
  if (k > k1) {
    k2 = k1;
    k1 = k;
  } else if (k > k2) {
    k2 = k;
  }
}

// Base pointer of a GridRef.  Written as selects on the (uniform) buffer id
// so the kernel argument block is never copied to scratch for indexing.
__device__ __forceinline__ float *base_of(const Bases &b, int buf) {
  float *p = buf == kBufMarginal ? b.marginal : b.scratch;
  p = buf == kBufUnary ? const_cast<float *>(b.unary) : p;
  p = buf == kBufRoot ? b.root : p;
  return p;
}
__device__ __forceinline__ const float *cptr(const Bases &b, const GridRef &g) { return base_of(b, g.buf) + g.off; }
__device__ __forceinline__ float *mptr(const Bases &b, const GridRef &g) { return base_of(b, g.buf) + g.off; }

__device__ __forceinline__ bool masked(const UnaryMask &m, int r, int x, int W) {
  return ((m.rot_bits >> r) & 1ull) || x < m.strip || x >= W - m.strip;
}

// log of a filter response, 0 -> LOG_ZERO (multi_array_op.hpp:154-166), on
// the hardware log2 (v_log_f32, about 1 ulp): a response the hardware takes
// for 0 (below 2^-126 where it flushes) is LOG_ZERO too
__device__ __forceinline__ float log0(float v) {
  const float l = __builtin_amdgcn_logf(v);
  return l > -INFINITY ? l * 0.693147180559945309f : kLogZero;
}

// The caller's accumulation of a column pass specialised per destination
// kind (EpiKind), so that a kernel's loads for it are unconditional
// straight-line code.
template <int EPI>
struct Epi {
  static constexpr bool kNeedU = EPI >= kEpiUpNodeU;  // reads the unary of the part written
  static constexpr bool kNeedB = EPI == kEpiDownMid;  // reads the stored upward belief
  static constexpr bool kDown = EPI >= kEpiDownLeaf;
  // grids at a second row pitch (byte offset ebs): the unary a chain-node
  // belief adds (the belief is the plan's grid, the unary the caller's), or
  // the stored belief and the next source of a five-grid downward epilogue
  // (the marginal and the unary it adds are at dst's pitch)
  static constexpr bool kSide = EPI == kEpiUpNodeU || EPI == kEpiDownMid;
  float vmax = -INFINITY;  // max of the written next source / belief
  unsigned best_key = 0u, best_eb = 0u;  // argmax of the marginal (key, byte offset in the slice)
  unsigned second_key = 0u;              // second largest marginal value (key) of the lane

  // one output at byte offset eb of slice r of dst, ebs of dst2 (both point
  // at the slice); pu / pb are only read when the kind needs them
  __device__ __forceinline__ void store(float *dst, float *dst2, unsigned eb, unsigned ebs, float msg, float pu,
                                        float pb, bool xmask, bool leaf_u) {
    auto at = [](float *p, unsigned o) -> float & { return *reinterpret_cast<float *>(reinterpret_cast<char *>(p) + o); };
    if (EPI == kEpiUpRoot || EPI == kEpiUpNode) {
      at(dst, eb) = msg;
      if (EPI == kEpiUpNode) vmax = fmaxf(vmax, msg);
    } else if (EPI == kEpiUpNodeU) {
      const float v = msg + (xmask ? kLogZero : pu);
      at(dst, eb) = v;
      vmax = fmaxf(vmax, v);
    } else {
      const float u = xmask ? kLogZero : pu;
      const float mg = (kNeedB ? pb : (leaf_u ? u : 0.0f)) + msg;
      at(dst, eb) = mg;
      const unsigned k = f2key(mg);
      if (k > best_key) {  // a lane's outputs come in ascending flat order: the first maximum stays
        second_key = best_key;
        best_key = k;
        best_eb = eb;
      } else if (k > second_key) {
        second_key = k;
      }
      if (kNeedB) {
        const float v2 = u + msg;
        at(dst2, ebs) = v2;
        vmax = fmaxf(vmax, v2);
      }
    }
  }
  // a padding column of the plan's grids (past W, inside the row pitch):
  // LOG_ZERO, so that every line of such a row is written whole by one wave
  // (no reduction sees it)
  __device__ __forceinline__ void pad(float *dst, float *dst2, unsigned eb, unsigned ebs, bool pad_a, bool pad_b) {
    auto at = [](float *p, unsigned o) -> float & { return *reinterpret_cast<float *>(reinterpret_cast<char *>(p) + o); };
    if (pad_a) at(dst, eb) = kLogZero;
    if (kNeedB && pad_b) at(dst2, ebs) = kLogZero;
  }
  // wave reductions and one atomic per wave; every lane of the wave calls it
  __device__ __forceinline__ void flush(const Bases &b, const MsgJob &J, long long slice0, int lane) {
    if (EPI == kEpiUpNode || EPI == kEpiUpNodeU || EPI == kEpiDownMid) {
      const unsigned k = wave_max_u32(vmax > -INFINITY ? f2key(vmax) : 0u);
      if (lane == 0 && k) atomicMax(b.max_keys + (kDown ? J.dst2_key : J.dst_key), k);
    }
    if (kDown) {
      const unsigned flat = (unsigned)(slice0 + (best_eb >> 2));
      const unsigned long long m =
          wave_max_u64(best_key ? ((unsigned long long)best_key << 32) | (0xffffffffull - flat) : 0ull);
      if (lane == 0 && m) atomicMax(b.map_keys + J.map_slot, m);
      if (b.top2_keys) top2_flush(b.top2_keys + J.map_slot, best_key, second_key, lane);
    }
  }
};

// Row pitch of an epilogue's side grids (Epi::kSide): the unary's for a
// chain-node belief, the next source's (and stored belief's) for the
// five-grid downward epilogue; dst's otherwise
template <int EPI>
__device__ __forceinline__ int epi_pitch_b(const MsgJob &J) {
  return EPI == kEpiUpNodeU ? J.dst_unary.pitch : (EPI == kEpiDownMid ? J.dst2.pitch : J.dst.pitch);
}

// Reductions a message's epilogue feeds: the max of the next upward source,
// the argmax of a marginal and the max of the next downward source.
struct DstAcc {
  unsigned kmax1 = 0u, kmax2 = 0u;
  unsigned long long amax = 0ull;
  unsigned second = 0u;  // second largest marginal value (key)
};

// The caller's accumulation of one message value msg at element
// flat = r*H*W + y*W + x of slice r, stored at dst[e] / dst2[e2] (each grid's
// own element of (r, y, x) at its row pitch).  pu is the unary of the part
// written (read when need_u), pb the stored upward belief (read when need_b),
// xmask the part's mask at (r, x).
//   up, parent = root:  the message slot (objectdetect_findrot.cpp:637-649)
//   up, chain node:     belief = 0 + message (+ unary) (:637, :652-654)
//   down:               marginal = belief + message (:201), next downward
//                       source = unary + message (:220-222)
__device__ __forceinline__ void dst_store(const MsgJob &J, float *dst, float *dst2, long long e, long long e2,
                                          unsigned flat, bool xmask, float msg, float pu, float pb, DstAcc &a) {
  if (J.dst_kind == kDstUpRoot) {
    dst[e] = msg;
  } else if (J.dst_kind == kDstUpNode) {
    float v = msg;
    if (J.dst_add_unary) v = v + (xmask ? kLogZero : pu);
    dst[e] = v;
    a.kmax1 = max(a.kmax1, f2key(v));
  } else {
    const float u = xmask ? kLogZero : pu;
    const float base = J.dst_base == 0 ? pb : (J.dst_base == 1 ? u : 0.0f);
    const float mg = base + msg;
    dst[e] = mg;
    const unsigned long long k = amax_key(mg, flat);
    const unsigned kv = (unsigned)(k >> 32), kb = (unsigned)(a.amax >> 32);
    a.second = kv > kb ? kb : max(a.second, kv);
    a.amax = k > a.amax ? k : a.amax;
    if (dst2) {
      const float v2 = u + msg;
      dst2[e2] = v2;
      a.kmax2 = max(a.kmax2, f2key(v2));
    }
  }
}

// Wave reductions of DstAcc and one atomic per wave; every lane of the wave
// calls it.
__device__ __forceinline__ void dst_flush(const Bases &b, const MsgJob &J, bool has_dst2, DstAcc &a, int lane) {
  if (J.dst_kind == kDstUpNode) {
    const unsigned v = wave_max_u32(a.kmax1);
    if (lane == 0 && v) atomicMax(b.max_keys + J.dst_key, v);
  } else if (J.dst_kind == kDstDown) {
    const unsigned long long m = wave_max_u64(a.amax);
    if (lane == 0 && m) atomicMax(b.map_keys + J.map_slot, m);
    if (b.top2_keys) top2_flush(b.top2_keys + J.map_slot, (unsigned)(a.amax >> 32), a.second, lane);
    if (has_dst2) {
      const unsigned v = wave_max_u32(a.kmax2);
      if (lane == 0 && v) atomicMax(b.max_keys + J.dst2_key, v);
    }
  }
}

}  // namespace ps
