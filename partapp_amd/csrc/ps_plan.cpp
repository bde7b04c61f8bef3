// Host-side planning: model validation, level schedule, message plans.
//
// The integer plan of a message reproduces the reference's double-precision
// coordinate arithmetic exactly (objectdetect_findrot.cpp:316-359, 437-448;
// homogeneous_coord.cpp:51-104; multi_array_transform.hpp:193-216), so the
// nearest-neighbour warps become one integer shift per rotation slice; a
// shift that is not uniform over the grid is reported as PS_ERR_WARP rather
// than silently approximated.
#include "ps_plan.hpp"

#include <cmath>
#include <cstdio>
#include <string>

#include "ps_internal.hpp"

namespace ps {

thread_local std::string g_last_error;

int set_error(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

// partapp_aux.hpp:44-55 with the float proto fields promoted to double
static double value_from_index(double minval, double maxval, double num_steps, int idx) {
  if (minval == maxval) return minval;
  double step_size = (maxval - minval) / num_steps;
  return minval + step_size * (0.5 + idx);
}

double rot_from_index(const ps_exp_param &ep, int rotidx) {
  return value_from_index(ep.min_part_rotation, ep.max_part_rotation, ep.num_rotation_steps, rotidx);
}
double scale_from_index(const ps_exp_param &ep, int scaleidx) {
  return value_from_index(ep.min_object_scale, ep.max_object_scale, ep.num_scale_steps, scaleidx);
}
// partapp_aux.hpp:25-42, -1 when out of range
int index_from_rot(const ps_exp_param &ep, double val) {
  double minval = ep.min_part_rotation, maxval = ep.max_part_rotation;
  if (minval == maxval) return 0;
  if (!(val >= minval && val < maxval)) return -1;
  double step_size = (maxval - minval) / (double)ep.num_rotation_steps;
  return (int)(unsigned)std::floor((val - minval) / step_size);
}

unsigned long long upright_masked_bits(const ps_exp_param &ep) {
  unsigned long long bits = 0;
  for (int r = 0; r < (int)ep.num_rotation_steps; ++r)
    if (!(std::fabs(rot_from_index(ep, r)) < 15.0)) bits |= 1ull << r;
  return bits;
}

// boost_math.cpp:104-117, unnormalised
static void gaussian_taps(double sigma, std::vector<double> &f) {
  int ksize = (int)std::floor(3 * sigma + 0.5);
  f.assign(2 * ksize + 1, 0.0);
  f[ksize] = 1.0;
  for (int i = 1; i <= ksize; ++i) {
    f[ksize + i] = std::exp(-i * i / (2 * sigma * sigma));
    f[ksize - i] = f[ksize + i];
  }
}

// Offset of the joint in the grid frame for rotation bin r: R(alpha) S(scale) v
// with alpha rounded to float (objectdetect_findrot.cpp:348-350, 439-441).
// uBLAS sums each product from zero; the zero terms of the rotation-scaling
// matrix leave the two products below unchanged.
static void rotated_offset(const ps_exp_param &ep, int r, double scale, const double v[2], double &dx, double &dy) {
  float alpha = rot_from_index(ep, r) * M_PI / 180.0;
  double ca = std::cos((double)alpha), sa = std::sin((double)alpha);
  double t00 = ca * scale, t01 = -sa * scale, t10 = sa * scale, t11 = ca * scale;
  double p0 = t00 * v[0];
  double p1 = t01 * v[1];
  dx = p0 + p1;
  double q0 = t10 * v[0];
  double q1 = t11 * v[1];
  dy = q0 + q1;
}

// One warp's integer offset along an axis of length n: the source index of
// output i is floor((i + shift) + 0.5) where shift = -d (forward warp, the
// inverse of a translation by +d) or +d (inverse warp).  Returns false when
// index - i is not the same for every i.
static bool uniform_offset(double shift, int n, int &off) {
  int o0 = 0;
  for (int i = 0; i < n; ++i) {
    double p = (double)i + shift;
    int o = (int)std::floor(p + 0.5) - i;
    if (i == 0) o0 = o;
    else if (o != o0) return false;
  }
  off = o0;
  return true;
}

int build_message_plan(const ps_exp_param &ep, int H, int W, const double offset_c[2], const double offset_p[2],
                       const double C[4], double rot_mean, double rot_sigma, double scale, MessagePlan &out) {
  const int R = (int)ep.num_rotation_steps;
  // objectdetect_findrot.cpp:316-323
  double rot_step_size = (ep.max_part_rotation - ep.min_part_rotation) / ep.num_rotation_steps;
  rot_step_size *= M_PI / 180.0;
  if (!(rot_step_size > 0)) return set_error(PS_ERR_INVALID_ARG, "rotation step must be > 0");
  double rot_sigma_idx = rot_sigma / rot_step_size;
  out.delta = (int)std::floor(-rot_mean / rot_step_size + 0.5);

  // rotation kernel (:377-420): rot_sigma > 0 filters, == 0 copies, and any
  // other value leaves the value-initialised (zero) result
  out.rot_taps.clear();
  if (rot_sigma > 0) {
    std::vector<double> f_rot;
    gaussian_taps(rot_sigma_idx, f_rot);
    int firstidx = 0;
    int f_rot_size = (int)f_rot.size();
    if (f_rot_size >= R) {
      int crot = f_rot_size / 2;
      f_rot_size = (R % 2 == 1) ? R - 2 : R - 1;
      firstidx = crot - f_rot_size / 2;
    }
    if (f_rot_size < 1 || f_rot_size % 2 != 1 || f_rot_size >= kMaxTaps)
      return set_error(PS_ERR_KERNEL_SIZE, "rotation kernel cannot be clipped to the rotation grid (R=" +
                                               std::to_string(R) + ")");
    for (int i = firstidx; i < firstidx + f_rot_size; ++i) out.rot_taps.push_back((float)f_rot[i]);
    out.rot_mode = kRotFilter;
  } else if (rot_sigma == 0) {
    out.rot_mode = kRotCopy;
  } else {
    out.rot_mode = kRotZero;
  }

  // spatial kernel: gaussFilter2d(scale^2 C) (:422-429, multi_array_filter.hpp:374-388)
  double sC[4];
  for (int i = 0; i < 4; ++i) sC[i] = (scale * scale) * C[i];
  if (!(sC[1] == 0 && sC[2] == 0))
    return set_error(PS_ERR_UNSUPPORTED, "off-diagonal covariance (gaussFilter2dOffset) is not supported");
  if (!(sC[0] > 0 && sC[3] > 0)) return set_error(PS_ERR_COVARIANCE, "covariance diagonal must be > 0");
  std::vector<double> fx, fy;
  gaussian_taps(std::sqrt(sC[0]), fx);
  gaussian_taps(std::sqrt(sC[3]), fy);
  if ((int)fx.size() >= kMaxTaps || (int)fy.size() >= kMaxTaps)
    return set_error(PS_ERR_KERNEL_SIZE, "spatial kernel must have fewer than 1000 taps");
  out.fx.assign(fx.begin(), fx.end());
  out.fy.assign(fy.begin(), fy.end());
  out.kx = ((int)fx.size() - 1) / 2;
  out.ky = ((int)fy.size() - 1) / 2;

  out.fwd_ox.assign(R, 0);
  out.fwd_oy.assign(R, 0);
  out.inv_ox.assign(R, 0);
  out.inv_oy.assign(R, 0);
  for (int r = 0; r < R; ++r) {
    double dx, dy;
    rotated_offset(ep, r, scale, offset_c, dx, dy);
    bool ok = uniform_offset(-dx, W, out.fwd_ox[r]) && uniform_offset(-dy, H, out.fwd_oy[r]);
    rotated_offset(ep, r, scale, offset_p, dx, dy);
    ok = ok && uniform_offset(dx, W, out.inv_ox[r]) && uniform_offset(dy, H, out.inv_oy[r]);
    if (!ok) return set_error(PS_ERR_WARP, "nearest-neighbour warp is not a uniform shift at rotation " + std::to_string(r));
  }
  return PS_OK;
}

int check_joints(const Model &m, const ps_joint *joints) {
  if (!joints) return set_error(PS_ERR_INVALID_ARG, "joints is NULL");
  for (int j = 0; j < m.P - 1; ++j)
    if (joints[j].child_idx != m.joints[j].child_idx || joints[j].parent_idx != m.joints[j].parent_idx)
      return set_error(PS_ERR_TREE, "instance joints must have the model's topology (joint " + std::to_string(j) + ")");
  return PS_OK;
}

int build_model(const ps_model_desc *d, Model &m) {
  if (!d) return set_error(PS_ERR_INVALID_ARG, "model descriptor is NULL");
  const int P = d->num_parts;
  if (P < 2 || P > 1024) return set_error(PS_ERR_INVALID_ARG, "num_parts must be in [2, 1024]");
  if (!d->is_detect || !d->is_upright || !d->is_root || !d->joints)
    return set_error(PS_ERR_INVALID_ARG, "null part flags or joints");
  const ps_exp_param &ep = d->exp_param;
  if (ep.num_rotation_steps < 1 || ep.num_rotation_steps > (unsigned)kMaxRot)
    return set_error(PS_ERR_UNSUPPORTED, "num_rotation_steps must be in [1, 64]");
  if (ep.num_scale_steps < 1) return set_error(PS_ERR_INVALID_ARG, "num_scale_steps must be >= 1");
  if (ep.min_object_scale == ep.max_object_scale && ep.num_scale_steps != 1)
    return set_error(PS_ERR_INVALID_ARG, "min_object_scale == max_object_scale needs num_scale_steps == 1");
  if (!(ep.max_part_rotation > ep.min_part_rotation))
    return set_error(PS_ERR_INVALID_ARG, "max_part_rotation must exceed min_part_rotation");
  if (ep.strip_border_detections > 0 && !(ep.strip_border_detections < 0.5))
    return set_error(PS_ERR_INVALID_ARG, "strip_border_detections must be < 0.5");

  m.P = P;
  m.ep = ep;
  m.is_sparse = d->is_sparse;
  m.is_detect.assign(d->is_detect, d->is_detect + P);
  m.is_upright.assign(d->is_upright, d->is_upright + P);
  Tree &t = m.tree;
  t.P = P;
  t.root = -1;
  for (int p = 0; p < P; ++p)
    if (d->is_detect[p] && d->is_root[p]) {
      if (t.root != -1) return set_error(PS_ERR_TREE, "more than one root part");
      t.root = p;
    }
  if (t.root < 0) return set_error(PS_ERR_TREE, "root part not found (is_detect && is_root)");
  if (d->num_joints != P - 1) return set_error(PS_ERR_TREE, "num_joints must be num_parts - 1");
  m.joints.assign(d->joints, d->joints + P - 1);

  t.parent_joint.assign(P, -1);
  t.child_joint.assign(P, -1);
  t.root_joints.clear();
  for (int j = 0; j < P - 1; ++j) {
    const ps_joint &J = m.joints[j];
    if (J.child_idx < 0 || J.child_idx >= P || J.parent_idx < 0 || J.parent_idx >= P || J.child_idx == J.parent_idx)
      return set_error(PS_ERR_TREE, "joint " + std::to_string(j) + " has an invalid part index");
    if (J.child_idx == t.root) return set_error(PS_ERR_TREE, "the root part cannot be a child");
    if (t.parent_joint[J.child_idx] != -1)
      return set_error(PS_ERR_TREE, "part " + std::to_string(J.child_idx) + " has two parents");
    t.parent_joint[J.child_idx] = j;
    if (J.parent_idx == t.root) {
      t.root_joints.push_back(j);
    } else {
      if (t.child_joint[J.parent_idx] != -1)
        return set_error(PS_ERR_TREE, "part " + std::to_string(J.parent_idx) +
                                          " has more than one child (only chains below the root are supported)");
      t.child_joint[J.parent_idx] = j;
    }
  }
  if ((int)t.root_joints.size() > kMaxRootChildren)
    return set_error(PS_ERR_UNSUPPORTED, "at most 16 children of the root");
  // depth, chain, acyclicity
  t.chain.assign(P, -1);
  t.depth.assign(P, 0);
  t.height.assign(P, 0);
  for (size_t k = 0; k < t.root_joints.size(); ++k) {
    int p = m.joints[t.root_joints[k]].child_idx;
    int d2 = 1;
    while (p >= 0) {
      if (t.chain[p] != -1 || d2 > P) return set_error(PS_ERR_TREE, "cycle in the joints");
      t.chain[p] = (int)k;
      t.depth[p] = d2++;
      int cj = t.child_joint[p];
      p = cj >= 0 ? m.joints[cj].child_idx : -1;
    }
  }
  for (int p = 0; p < P; ++p)
    if (p != t.root && t.chain[p] < 0) return set_error(PS_ERR_TREE, "part " + std::to_string(p) + " is not connected to the root");
  t.max_height = t.max_depth = 0;
  for (int p = 0; p < P; ++p) {
    if (p == t.root) continue;
    int h = 1, q = p;
    while (t.child_joint[q] >= 0) {
      q = m.joints[t.child_joint[q]].child_idx;
      ++h;
    }
    t.height[p] = h;
    t.max_height = std::max(t.max_height, h);
    t.max_depth = std::max(t.max_depth, t.depth[p]);
  }
  return PS_OK;
}

}  // namespace ps
