// ps_message_noalloc: one message (computeRotJointMarginal) that allocates no
// device memory.  The block ps_message allocates per call (tables, keys, the
// row-pass output grid, eigenbasis grids) lives here in a file-scope device
// array instead, `scratch`, of PS_MESSAGE_NOALLOC_BYTES bytes: every GPU the
// library runs on holds one copy of it from module load to process end.
//
// Host code never passes `scratch` itself to a kernel or a copy (its name on
// the host is a stand-in, not a device address): the first call on a device
// looks up that device's address of the array with hipGetSymbolAddress, with
// the device current, and every later call on it reuses the stored address.
// All calls on one GPU share its array, which keeps the last call's contents;
// the kernels clear the two keys they accumulate into before using them, and
// every other byte they read was written earlier in the same call.  Calls
// are serialised by a process-wide lock (each call is synchronous), so calls
// from several host threads do not overlap on the array; two processes each
// have their own copy.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ps_message.hpp"

static __device__ __align__(256) unsigned char scratch[PS_MESSAGE_NOALLOC_BYTES];

namespace {
std::mutex g_mu;                 // guards g_addr and the arrays' contents
std::vector<char *> g_addr;      // device address of `scratch` per device (null: not looked up yet)
}  // namespace

using namespace ps;

extern "C" int ps_message_noalloc(const ps_model *model, const float *child_dev, float *parent_dev, int height,
                                  int width, const double offset_c[2], const double offset_p[2], const double C[4],
                                  double rot_mean, double rot_sigma, double scale, int is_sparse, void *stream) {
  if (!child_dev || !parent_dev) return set_error(PS_ERR_INVALID_ARG, "null argument");
  MessageSetup ms;
  int st = message_prepare(model, height, width, offset_c, offset_p, C, rot_mean, rot_sigma, scale, is_sparse, ms);
  if (st != PS_OK) return st;
  if (ms.total > (size_t)PS_MESSAGE_NOALLOC_BYTES)
    return set_error(PS_ERR_UNSUPPORTED, "message needs " + std::to_string(ms.total) +
                                             " bytes of scratch; ps_message_noalloc has " +
                                             std::to_string((size_t)PS_MESSAGE_NOALLOC_BYTES) +
                                             " (use ps_message for this size)");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(PS_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_mu);
  if ((int)g_addr.size() <= dev) g_addr.resize(dev + 1, nullptr);
  if (!g_addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(PS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_addr[dev] = static_cast<char *>(p);
  }
  return message_enqueue(ms, g_addr[dev], child_dev, parent_dev, static_cast<hipStream_t>(stream));
}
