// One message (computeRotJointMarginal) run on its own, outside a plan: the
// tables, keys and scratch grids of the message laid out in one block of
// device memory (ps_message allocates the block per call,
// ps_message_noalloc uses a file-scope device array).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "../../include/partapp_ps.h"

namespace ps {

struct MessageSetup {
  int R = 0, H = 0, W = 0;
  bool od = false;          // off-diagonal covariance: eigenbasis stages instead of the column pass
  bool split = false;       // a warp whose offset changes along an axis (MsgJob::split)
  long long od_cells = 0;   // R * th * tw of the eigenbasis grid
  size_t lds = 0, col_lds = 0;
  // byte offsets in the block: job, max job, taps, shifts, max key, argmax key, grids
  size_t o_job = 0, o_max = 0, o_taps = 0, o_sh = 0, o_key = 0, o_mk = 0, o_T = 0;
  size_t total = 0;         // bytes of the block
  std::vector<char> host;   // the first o_T bytes of the block (the tables)
  int epi = 0;              // epilogue kind of the column pass
};

// Validate the arguments and build the message's tables and layout.
int message_prepare(const ps_model *model, int height, int width, const double offset_c[2], const double offset_p[2],
                    const double C[4], double rot_mean, double rot_sigma, double scale, int is_sparse,
                    MessageSetup &ms);

// Upload the tables into `mem` (at least ms.total bytes of device memory on
// the current device), run the message on stream s and wait for it.
int message_enqueue(const MessageSetup &ms, char *mem, const float *child_dev, float *parent_dev, hipStream_t s);

int set_error(int status, const std::string &msg);

}  // namespace ps
