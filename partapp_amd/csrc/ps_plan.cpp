// Host-side planning: model validation, level schedule, message plans.
//
// The integer plan of a message reproduces the reference's double-precision
// coordinate arithmetic exactly (objectdetect_findrot.cpp:316-359, 437-448;
// homogeneous_coord.cpp:51-104; multi_array_transform.hpp:193-216), so the
// nearest-neighbour warps become one integer shift per rotation slice and
// axis, or two where the double sum's rounding changes it once along the
// axis (a shift within an ulp of a half-integer); a shift that changes more
// than once is reported as PS_ERR_WARP rather than silently approximated.
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
// floor(i + shift + 0.5) - i for i in [0, n) (multi_array_transform.hpp:
// 197-205 on a translation): off below index brk, off + d from brk on
// (brk = n, d = 0 when it is one value); false if it changes more than once
static bool axis_offset(double shift, int n, int &off, int &brk, int &d) {
  off = (int)std::floor(shift + 0.5);
  brk = n;
  d = 0;
  for (int i = 1; i < n; ++i) {
    const double p = (double)i + shift;
    const int o = (int)std::floor(p + 0.5) - i;
    if (o == (i < brk ? off : off + d)) continue;
    if (brk != n) return false;
    brk = i;
    d = o - off;
  }
  return true;
}

// ---- 3x3 homogeneous matrices for the eigenbasis transforms, row major.
// Products sum from zero in index order and the inverse follows
// homogeneous_coord.cpp:51-71, so the coordinates match the reference's
// double arithmetic.
namespace {
struct Mat3 {
  double m[3][3];
};
Mat3 mat_zero() {
  Mat3 r;
  for (auto &row : r.m)
    for (double &v : row) v = 0.0;
  return r;
}
Mat3 mat_ident() {
  Mat3 r = mat_zero();
  r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0;
  return r;
}
Mat3 mat_mul(const Mat3 &a, const Mat3 &b) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += a.m[i][k] * b.m[k][j];
      r.m[i][j] = t;
    }
  return r;
}
Mat3 mat_inverse(const Mat3 &T) {
  Mat3 r;
  const double D = T.m[0][0] * T.m[1][1] - T.m[1][0] * T.m[0][1];
  r.m[0][0] = T.m[1][1] / D;
  r.m[0][1] = -T.m[0][1] / D;
  r.m[1][0] = -T.m[1][0] / D;
  r.m[1][1] = T.m[0][0] / D;
  for (int i = 0; i < 2; ++i) {
    double t = 0.0;
    for (int k = 0; k < 2; ++k) t += r.m[i][k] * T.m[k][2];
    r.m[i][2] = -t;
  }
  r.m[2][0] = 0;
  r.m[2][1] = 0;
  r.m[2][2] = 1;
  return r;
}
Mat3 mat_homogeneous(double r00, double r01, double r10, double r11, double dx, double dy) {
  Mat3 T = mat_zero();
  T.m[0][0] = r00;
  T.m[0][1] = r01;
  T.m[1][0] = r10;
  T.m[1][1] = r11;
  T.m[0][2] = dx;
  T.m[1][2] = dy;
  T.m[2][2] = 1;
  return T;
}
Mat3 mat_translation(double dx, double dy) { return mat_homogeneous(1, 0, 0, 1, dx, dy); }
void mat_store(const Mat3 &M, double out[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = M.m[i][j];
}

// boost_math.cpp:40-99: eigenvalues ascending; V's first column belongs to
// the smallest, V = [v1 v2] with v2 = (-v21, v11)
void eigen_2x2(const double M[4], double V[4], double E[2]) {
  const double m11 = M[0], m12 = M[1], m22 = M[3];
  double e1, e2, v11, v21;
  if (m12 != 0) {
    const double sqrtD = std::sqrt((m11 - m22) * (m11 - m22) + 4 * m12 * m12);
    e1 = 0.5 * (m11 + m22 - sqrtD);
    e2 = 0.5 * (m11 + m22 + sqrtD);
    v11 = 0.5 * (m11 - m22 - sqrtD) / m12;
    v21 = 1;
  } else if (m11 < m22) {
    e1 = m11;
    e2 = m22;
    v11 = 1;
    v21 = 0;
  } else {
    e1 = m22;
    e2 = m11;
    v11 = 0;
    v21 = 1;
  }
  const double norm_v1 = std::sqrt(v11 * v11 + v21 * v21);
  v11 /= norm_v1;
  v21 /= norm_v1;
  V[0] = v11;
  V[1] = -v21;
  V[2] = v21;
  V[3] = v11;
  E[0] = e1;
  E[1] = e2;
}
}  // namespace

void warp_matrix(const double T21[9], bool gather, double out[6]) {
  Mat3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.m[i][j] = T21[3 * i + j];
  const Mat3 M = gather ? mat_mul(mat_inverse(T), mat_ident()) : mat_mul(mat_ident(), T);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = M.m[i][j];
}

// The eigenbasis path of one message: multi_array_filter.hpp:334-369 with
// offset 0, transform_grid_resize (multi_array_transform.hpp:284-305) and
// get_transformed_bbox (homogeneous_coord.cpp:128-146).
static int plan_offdiagonal(const double sC[4], int H, int W, bool is_sparse, MessagePlan &out) {
  double V[4], E[2];
  eigen_2x2(sC, V, E);
  if (!(E[0] > 0 && E[1] > 0)) return set_error(PS_ERR_COVARIANCE, "covariance must be positive definite");
  const Mat3 T21 = mat_homogeneous(V[0], V[2], V[1], V[3], 0, 0);  // V transposed
  const double cx[4] = {0.0, (double)(W - 1), 0.0, (double)(W - 1)};
  const double cy[4] = {0.0, 0.0, (double)(H - 1), (double)(H - 1)};
  double minx = 0, maxx = 0, miny = 0, maxy = 0;
  for (int i = 0; i < 4; ++i) {
    const double p[3] = {cx[i], cy[i], 1.0};
    double q[2];
    for (int r = 0; r < 2; ++r) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += T21.m[r][k] * p[k];
      q[r] = t;
    }
    if (i == 0 || q[0] < minx) minx = q[0];
    if (i == 0 || q[0] > maxx) maxx = q[0];
    if (i == 0 || q[1] < miny) miny = q[1];
    if (i == 0 || q[1] > maxy) maxy = q[1];
  }
  out.tw = (int)std::ceil(maxx - minx);
  out.th = (int)std::ceil(maxy - miny);
  if (out.tw < 1 || out.th < 1) return set_error(PS_ERR_INVALID_ARG, "degenerate eigenbasis grid");
  const Mat3 T23 = mat_translation(minx, miny), T32 = mat_translation(-minx, -miny);
  out.fwd_direct = is_sparse;
  if (is_sparse)
    mat_store(mat_mul(T32, T21), out.t_fwd);
  else
    mat_store(mat_mul(mat_inverse(T21), T23), out.t_fwd);
  const Mat3 T42 = mat_homogeneous(V[0], V[1], V[2], V[3], -0.0, -0.0);
  mat_store(mat_mul(mat_inverse(mat_mul(T42, T23)), mat_ident()), out.t_back);
  std::vector<double> fx, fy;
  gaussian_taps(std::sqrt(E[0]), fx);
  gaussian_taps(std::sqrt(E[1]), fy);
  if ((int)fx.size() >= kMaxTaps || (int)fy.size() >= kMaxTaps)
    return set_error(PS_ERR_KERNEL_SIZE, "spatial kernel must have fewer than 1000 taps");
  out.fx.assign(fx.begin(), fx.end());
  out.fy.assign(fy.begin(), fy.end());
  out.kx = ((int)fx.size() - 1) / 2;
  out.ky = ((int)fy.size() - 1) / 2;
  out.od = true;
  return PS_OK;
}

int build_message_plan(const ps_exp_param &ep, int H, int W, const double offset_c[2], const double offset_p[2],
                       const double C[4], double rot_mean, double rot_sigma, double scale, bool is_sparse,
                       MessagePlan &out) {
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
  out.od = false;
  if (!(sC[1] == 0 && sC[2] == 0)) {
    int st = plan_offdiagonal(sC, H, W, is_sparse, out);
    if (st != PS_OK) return st;
  } else {
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
  }

  for (auto *v : {&out.fwd_ox, &out.fwd_oy, &out.inv_ox, &out.inv_oy, &out.fwd_xd, &out.fwd_yd, &out.inv_xd,
                  &out.inv_yd})
    v->assign(R, 0);
  out.fwd_xb.assign(R, W);
  out.inv_xb.assign(R, W);
  out.fwd_yb.assign(R, H);
  out.inv_yb.assign(R, H);
  for (int r = 0; r < R; ++r) {
    double dx, dy;
    rotated_offset(ep, r, scale, offset_c, dx, dy);
    bool ok = axis_offset(-dx, W, out.fwd_ox[r], out.fwd_xb[r], out.fwd_xd[r]) &&
              axis_offset(-dy, H, out.fwd_oy[r], out.fwd_yb[r], out.fwd_yd[r]);
    rotated_offset(ep, r, scale, offset_p, dx, dy);
    ok = ok && axis_offset(dx, W, out.inv_ox[r], out.inv_xb[r], out.inv_xd[r]) &&
         axis_offset(dy, H, out.inv_oy[r], out.inv_yb[r], out.inv_yd[r]);
    if (!ok)
      return set_error(PS_ERR_WARP, "nearest-neighbour warp changes its offset more than once at rotation " +
                                        std::to_string(r));
    if (out.fwd_xb[r] < W || out.fwd_yb[r] < H || out.inv_xb[r] < W || out.inv_yb[r] < H) out.split = true;
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
