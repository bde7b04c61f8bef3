// Host-side planning of the sweep: model validation, the level schedule of
// the tree, and the integer plan + filter taps of every message.
#pragma once

#include <string>
#include <vector>

#include "../../include/partapp_ps.h"

namespace ps {

// Tree of one model: one root and chains hanging off it
// (objectdetect_findrot.cpp:207-210 requires every non-root part to have at
// most one child).
struct Tree {
  int P = 0;
  int root = -1;
  std::vector<int> parent_joint;  // [P] joint whose child is p, -1 for the root
  std::vector<int> child_joint;   // [P] joint whose parent is p (non-root), -1 for leaves
  std::vector<int> root_joints;   // joints with parent == root, in joint order
  std::vector<int> chain;         // [P] index of the root child whose chain holds p, -1 for root
  std::vector<int> height;        // [P] 1 for leaves, 1 + height(child) above
  std::vector<int> depth;         // [P] 0 for the root, 1 for its children, ...
  int max_height = 0;             // number of upward levels
  int max_depth = 0;              // number of downward levels
};

struct Model {
  int P = 0;
  std::vector<int> is_detect, is_upright;
  std::vector<ps_joint> joints;
  ps_exp_param ep{};
  int is_sparse = 0;
  Tree tree;
};

// Integer plan and taps of one message (computeRotJointMarginal,
// objectdetect_findrot.cpp:292-456) for a given grid size and scale.
struct MessagePlan {
  int delta = 0;                     // rot_mean_idx (:323)
  int rot_mode = 0;                  // RotMode
  std::vector<float> rot_taps;       // clipped circular kernel (:377-404)
  std::vector<float> fx, fy;         // spatial taps, float (multi_array_filter.hpp:246-250)
  int kx = 0, ky = 0;
  std::vector<int> fwd_ox, fwd_oy;   // [R] by source slice
  std::vector<int> inv_ox, inv_oy;   // [R] by slice
  // A nearest-neighbour warp whose offset changes once along an axis (the
  // double sum i + shift rounds differently in different binades of its
  // magnitude when shift lies within an ulp of a half-integer): from index
  // *_b on, the offset is the one above plus *_d.  b = the axis length, d = 0:
  // uniform.  split: some slice's warp changes along some axis.
  std::vector<int> fwd_xb, fwd_xd, fwd_yb, fwd_yd;  // [R] by source slice
  std::vector<int> inv_xb, inv_xd, inv_yb, inv_yd;  // [R] by slice
  bool split = false;
  // off-diagonal covariance (gaussFilter2dOffset, multi_array_filter.hpp:334-369):
  // rotate into the eigenbasis of scale^2 C onto a th x tw grid, filter
  // there with fx, fy (kx, ky above are the eigenbasis half widths), sample back
  bool od = false;
  int th = 0, tw = 0;
  bool fwd_direct = false;  // TM_DIRECT scatter (sparse) or TM_BILINEAR gather
  double t_fwd[9] = {};     // direct: source -> eigenbasis grid; bilinear: grid -> source
  double t_back[9] = {};    // output pixel -> sample position in the eigenbasis grid
};

// Status + message helpers; every failure path sets the thread's error text.
int set_error(int status, const std::string &msg);

int build_model(const ps_model_desc *desc, Model &m);
// joints must be P-1 joints with the model's topology
int check_joints(const Model &m, const ps_joint *joints);
int build_message_plan(const ps_exp_param &ep, int H, int W, const double offset_c[2], const double offset_p[2],
                       const double C[4], double rot_mean, double rot_sigma, double scale, bool is_sparse,
                       MessagePlan &out);

// Rows 0-1 of the sampling matrix of a fixed-size warp with the homogeneous
// 3x3 T21 (row major): T21 itself for TM_DIRECT (source -> output), else
// inverse(T21) for TM_BILINEAR / TM_NEAREST (output -> source)
// (multi_array_transform.hpp:131-136, 167-196).
void warp_matrix(const double T21[9], bool gather, double out[6]);

double rot_from_index(const ps_exp_param &ep, int rotidx);
double scale_from_index(const ps_exp_param &ep, int scaleidx);
int index_from_rot(const ps_exp_param &ep, double rot);
// rotations an upright part keeps: |rot| < 15 degrees (objectdetect_findrot.cpp:509-523)
unsigned long long upright_masked_bits(const ps_exp_param &ep);

}  // namespace ps
