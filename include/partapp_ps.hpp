// partapp_ps.hpp — C++ host interface with the reference's signatures.
//
// object_detect::computeRootPosteriorRot (objectdetect_findrot.cpp:470-727,
// objectdetect.h:269-274) and computeRotJointMarginal (:292-456,
// objectdetect.h:277-282) re-declared as templates over the caller's own
// types, implemented on the C ABI of partapp_ps.h (HIP device memory is
// allocated and freed per call, like the reference's per-call grids).
//
// The templates read only what the reference reads:
//   PartAppT   .m_exp_param.{min,max}_part_rotation(), num_rotation_steps(),
//              {min,max}_object_scale(), num_scale_steps(),
//              strip_border_detections(), roi_save_num_samples()
//              .m_part_conf.part_size(), .part(p).is_detect() / is_upright() / is_root()
//   JointT     child_idx, parent_idx, offset_c(i), offset_p(i), C(i,j), rot_mean, rot_sigma
//   Grid3      data(), shape()[0..2] (boost::multi_array<float,3> qualifies)
//   PartHypT   constructible as PartHypT(exp_param, scaleidx, rotidx, x, y, score)
// so partapp's PartApp, Joint, FloatGrid3 and PartHyp bind unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "partapp_ps.h"

namespace partapp_ps {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &where)
      : std::runtime_error(where + " failed (" + std::to_string(status) + "): " + ps_last_error_string()),
        status(status) {}
  int status;
};

inline void check(int st, const char *where) {
  if (st != PS_OK) throw Error(st, where);
}
inline void check_hip(hipError_t e, const char *where) {
  if (e != hipSuccess) throw std::runtime_error(std::string(where) + ": " + hipGetErrorString(e));
}

// RAII device buffer
struct DeviceBuffer {
  void *p = nullptr;
  explicit DeviceBuffer(size_t bytes) { check_hip(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc"); }
  ~DeviceBuffer() { (void)hipFree(p); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  float *f() const { return static_cast<float *>(p); }
};

template <class ExpParamT>
ps_exp_param to_exp_param(const ExpParamT &e) {
  ps_exp_param p;
  p.min_part_rotation = e.min_part_rotation();
  p.max_part_rotation = e.max_part_rotation();
  p.num_rotation_steps = e.num_rotation_steps();
  p.min_object_scale = e.min_object_scale();
  p.max_object_scale = e.max_object_scale();
  p.num_scale_steps = e.num_scale_steps();
  p.strip_border_detections = e.strip_border_detections();
  return p;
}

template <class JointT>
ps_joint to_joint(const JointT &j) {
  ps_joint o;
  o.child_idx = j.child_idx;
  o.parent_idx = j.parent_idx;
  for (int i = 0; i < 2; ++i) {
    o.offset_c[i] = j.offset_c(i);
    o.offset_p[i] = j.offset_p(i);
  }
  o.C[0] = j.C(0, 0);
  o.C[1] = j.C(0, 1);
  o.C[2] = j.C(1, 0);
  o.C[3] = j.C(1, 1);
  o.rot_mean = j.rot_mean;
  o.rot_sigma = j.rot_sigma;
  return o;
}

// Owns a ps_model built from the caller's PartApp and joints.
class Model {
 public:
  template <class PartAppT, class JointT>
  Model(const PartAppT &part_app, int rootpart_idx, const std::vector<JointT> &joints, bool bIsSparse) {
    const int P = part_app.m_part_conf.part_size();
    det_.resize(P);
    up_.resize(P);
    root_.assign(P, 0);
    for (int p = 0; p < P; ++p) {
      det_[p] = part_app.m_part_conf.part(p).is_detect() ? 1 : 0;
      up_[p] = part_app.m_part_conf.part(p).is_upright() ? 1 : 0;
    }
    root_[rootpart_idx] = 1;  // the caller chose the root (objectdetect_findrot.cpp:781-786)
    for (const JointT &j : joints) joints_.push_back(to_joint(j));
    ps_model_desc d;
    d.num_parts = P;
    d.is_detect = det_.data();
    d.is_upright = up_.data();
    d.is_root = root_.data();
    d.num_joints = (int)joints_.size();
    d.joints = joints_.data();
    d.exp_param = to_exp_param(part_app.m_exp_param);
    d.is_sparse = bIsSparse ? 1 : 0;
    check(ps_model_create(&d, &m_), "ps_model_create");
  }
  ~Model() { ps_model_destroy(m_); }
  Model(const Model &) = delete;
  Model &operator=(const Model &) = delete;
  ps_model *get() const { return m_; }

 private:
  std::vector<int> det_, up_, root_;
  std::vector<ps_joint> joints_;
  ps_model *m_ = nullptr;
};

// computeRootPosteriorRot with the reference's parameters.  Differences that
// the reference's callers cannot observe: the upright / border masks are
// written into log_part_detections exactly as the reference does
// (objectdetect_findrot.cpp:509-551) but computed on the GPU from an
// unmasked copy; bSaveMarginals asks for log_part_posterior_out instead of
// .mat files, and imgidx / flip only name those files in the reference.
template <class PartAppT, class Grid3, class JointT, class PartHypT>
void computeRootPosteriorRot(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                             Grid3 &root_part_posterior, int rootpart_idx, std::vector<JointT> joints, bool flip,
                             bool bIsSparse, int imgidx, std::vector<std::vector<PartHypT>> &best_part_hyp,
                             bool bSaveMarginals, std::vector<std::vector<Grid3>> *log_part_posterior_out = nullptr) {
  (void)flip;
  (void)imgidx;
  (void)bSaveMarginals;
  const auto &ep = part_app.m_exp_param;
  const int P = part_app.m_part_conf.part_size();
  const int S = (int)ep.num_scale_steps();
  const int R = (int)ep.num_rotation_steps();
  if ((int)log_part_detections.size() != P || (int)log_part_detections[0].size() != S)
    throw std::invalid_argument("log_part_detections must be [nParts][nScales]");
  const int H = (int)log_part_detections[0][0].shape()[1];
  const int W = (int)log_part_detections[0][0].shape()[2];
  const size_t n = (size_t)R * H * W;
  for (int p = 0; p < P; ++p)
    for (int s = 0; s < S; ++s) {
      const auto &g = log_part_detections[p][s];
      if ((int)g.shape()[0] != R || (int)g.shape()[1] != H || (int)g.shape()[2] != W)
        throw std::invalid_argument("every score grid must be [nRotations][H][W]");
    }

  Model model(part_app, rootpart_idx, joints, bIsSparse);
  std::vector<float> host(n * P * S);
  for (int p = 0; p < P; ++p)
    for (int s = 0; s < S; ++s) std::copy(log_part_detections[p][s].data(), log_part_detections[p][s].data() + n,
                                          host.begin() + ((size_t)p * S + s) * n);
  DeviceBuffer un(host.size() * sizeof(float)), mg(host.size() * sizeof(float)),
      rp((size_t)S * H * W * sizeof(float));
  check_hip(hipMemcpy(un.p, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  ps_instance inst;
  inst.height = H;
  inst.width = W;
  inst.unary_offset = inst.marginal_offset = inst.root_offset = 0;
  inst.joints = nullptr;
  std::vector<ps_map> maps((size_t)S * P);
  check(ps_marginals(model.get(), &inst, 1, un.f(), mg.f(), rp.f(), maps.data(), nullptr), "ps_marginals");

  std::vector<float> root((size_t)S * H * W);
  check_hip(hipMemcpy(root.data(), rp.p, root.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
  root_part_posterior.resize(std::array<size_t, 3>{(size_t)S, (size_t)H, (size_t)W});
  std::copy(root.begin(), root.end(), root_part_posterior.data());

  if (log_part_posterior_out) {
    check_hip(hipMemcpy(host.data(), mg.p, host.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    log_part_posterior_out->assign(P, std::vector<Grid3>(S, log_part_detections[0][0]));
    for (int p = 0; p < P; ++p)
      for (int s = 0; s < S; ++s)
        std::copy(host.begin() + ((size_t)p * S + s) * n, host.begin() + ((size_t)p * S + s + 1) * n,
                  (*log_part_posterior_out)[p][s].data());
  }

  // MAP of the last scale, then its local maxima (objectdetect_findrot.cpp:257-283
  // run per scale; findLocalMax tags its hypotheses with scale index 0,
  // objectdetect_aux.cpp:304)
  best_part_hyp.clear();
  best_part_hyp.resize(P);
  const int max_n = (int)ep.roi_save_num_samples();
  std::vector<ps_local_max> lm;
  std::vector<int> lm_n(P, 0);
  if (max_n > 0) {
    std::vector<long long> offs(P);
    std::vector<int> dims(3 * P);
    for (int p = 0; p < P; ++p) {
      offs[p] = ((long long)p * S + (S - 1)) * (long long)n;
      dims[3 * p] = R;
      dims[3 * p + 1] = H;
      dims[3 * p + 2] = W;
    }
    DeviceBuffer lmd(sizeof(ps_local_max) * (size_t)P * max_n), cnd(sizeof(int) * P);
    check(ps_find_local_max(mg.f(), offs.data(), dims.data(), P, max_n, static_cast<ps_local_max *>(lmd.p),
                            static_cast<int *>(cnd.p), nullptr),
          "ps_find_local_max");
    lm.resize((size_t)P * max_n);
    check_hip(hipMemcpy(lm.data(), lmd.p, lm.size() * sizeof(ps_local_max), hipMemcpyDeviceToHost), "hipMemcpy");
    check_hip(hipMemcpy(lm_n.data(), cnd.p, sizeof(int) * P, hipMemcpyDeviceToHost), "hipMemcpy");
  }
  for (int p = 0; p < P; ++p) {
    const ps_map &m = maps[(size_t)(S - 1) * P + p];
    best_part_hyp[p].push_back(PartHypT(ep, S - 1, m.rotidx, m.x, m.y, m.score));
    for (int i = 0; i < lm_n[p]; ++i) {
      const ps_local_max &h = lm[(size_t)p * max_n + i];
      best_part_hyp[p].push_back(PartHypT(ep, 0, h.d, h.x, h.y, h.score));
    }
  }

  // the reference's in-place masking of the caller's grids (:509-551)
  for (int p = 0; p < P; ++p) {
    if (!part_app.m_part_conf.part(p).is_upright()) continue;
    for (int s = 0; s < S; ++s)
      for (int r = 0; r < R; ++r) {
        double step = ((double)ep.max_part_rotation() - (double)ep.min_part_rotation()) / R;
        double rot = ep.min_part_rotation() == ep.max_part_rotation()
                         ? (double)ep.min_part_rotation()
                         : (double)ep.min_part_rotation() + step * (0.5 + r);
        if (!(std::fabs(rot) < 15.0)) {
          float *g = log_part_detections[p][s].data() + (size_t)r * H * W;
          std::fill(g, g + (size_t)H * W, -1e6f);
        }
      }
  }
  if (ep.strip_border_detections() > 0) {
    const int strip = (int)(ep.strip_border_detections() * W);
    for (int s = 0; s < S; ++s)
      for (size_t row = 0; row < (size_t)R * H; ++row) {
        float *g = log_part_detections[rootpart_idx][s].data() + row * W;
        for (int x = 0; x < strip; ++x) g[x] = -1e6f;
        for (int x = W - strip; x < W; ++x) g[x] = -1e6f;
      }
  }
}

// findLocalMax of a [nScales][H][W] root posterior into a HypothesisList
// (objectdetect_aux.cpp:266-286): one hypothesis per local maximum, appended,
// with scale = scale_from_index(scale index), x, y and score.  HypListT needs
// add_hyp() returning a pointer with set_scale / set_x / set_y / set_score
// (the protobuf HypothesisList of HypothesisList.proto qualifies).
template <class ExpParamT, class Grid3, class HypListT>
void findLocalMax(const ExpParamT &exp_param, const Grid3 &log_prob_grid, HypListT &hypothesis_list,
                  int max_hypothesis_number) {
  const int D = (int)log_prob_grid.shape()[0], H = (int)log_prob_grid.shape()[1], W = (int)log_prob_grid.shape()[2];
  if (max_hypothesis_number < 1) return;
  const size_t n = (size_t)D * H * W;
  DeviceBuffer g(n * sizeof(float)), lmd(sizeof(ps_local_max) * (size_t)max_hypothesis_number), cnd(sizeof(int));
  check_hip(hipMemcpy(g.p, log_prob_grid.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  long long off = 0;
  int dims[3] = {D, H, W};
  check(ps_find_local_max(g.f(), &off, dims, 1, max_hypothesis_number, static_cast<ps_local_max *>(lmd.p),
                          static_cast<int *>(cnd.p), nullptr),
        "ps_find_local_max");
  int count = 0;
  check_hip(hipMemcpy(&count, cnd.p, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  std::vector<ps_local_max> lm((size_t)count);
  if (count)
    check_hip(hipMemcpy(lm.data(), lmd.p, sizeof(ps_local_max) * (size_t)count, hipMemcpyDeviceToHost), "hipMemcpy");
  const ps_exp_param ep = to_exp_param(exp_param);
  for (const ps_local_max &h : lm) {
    // scale_from_index (partapp_aux.hpp:44-55) on the float proto fields
    const double mn = ep.min_object_scale, mx = ep.max_object_scale;
    const double scale = mn == mx ? mn : mn + (mx - mn) / (double)ep.num_scale_steps * (0.5 + h.d);
    auto *hyp = hypothesis_list.add_hyp();
    hyp->set_scale((float)scale);
    hyp->set_x((float)h.x);
    hyp->set_y((float)h.y);
    hyp->set_score(h.score);
  }
}

// The object hypotheses findObjectImageRotJoints derives from the root
// posterior (objectdetect_findrot.cpp:1037-1048): up to 1000 local maxima,
// each tagged with flip and moved from the root position to the bounding
// box by the window's bbox offset (truncated to int).  PartAppT also needs
// m_window_param.bbox_offset_x() / bbox_offset_y(); HypListT also needs
// hyp_size(), hyp(i).x() / y() and mutable_hyp(i)->set_flip / set_x / set_y.
template <class PartAppT, class Grid3, class HypListT>
void rootHypotheses(const PartAppT &part_app, const Grid3 &root_part_posterior, bool flip,
                    HypListT &hypothesis_list) {
  const int max_hypothesis_number = 1000;
  findLocalMax(part_app.m_exp_param, root_part_posterior, hypothesis_list, max_hypothesis_number);
  for (int i = 0; i < hypothesis_list.hyp_size(); ++i) {
    hypothesis_list.mutable_hyp(i)->set_flip(flip);
    const int bbox_x = (int)(hypothesis_list.hyp(i).x() + part_app.m_window_param.bbox_offset_x());
    const int bbox_y = (int)(hypothesis_list.hyp(i).y() + part_app.m_window_param.bbox_offset_y());
    hypothesis_list.mutable_hyp(i)->set_x(bbox_x);
    hypothesis_list.mutable_hyp(i)->set_y(bbox_y);
  }
}

// best_conf of findObjectImageRotJoints (objectdetect_findrot.cpp:1005-1008):
// [nParts][7], row p = best_part_hyp[p][0].toVect() = (scaleidx, scale,
// rotidx, rot, x, y, score) (objectdetect.h:139-159).
template <class PartHypT>
std::vector<float> bestConf(const std::vector<std::vector<PartHypT>> &best_part_hyp) {
  std::vector<float> out;
  out.reserve(best_part_hyp.size() * 7);
  for (const auto &hyps : best_part_hyp) {
    if (hyps.empty()) throw std::invalid_argument("best_part_hyp has a part without hypotheses");
    const auto v = hyps[0].toVect();
    for (int i = 0; i < 7; ++i) out.push_back((float)v[i]);
  }
  return out;
}

// MAP record of the convenience overload, which discards it.
struct IgnoredHyp {
  template <class E>
  IgnoredHyp(const E &, int, int, int, int, float) {}
};

// The two-output convenience overload (objectdetect_findrot.cpp:458-468).
template <class PartAppT, class Grid3, class JointT>
void computeRootPosteriorRot(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                             Grid3 &root_part_posterior, int rootpart_idx, std::vector<JointT> joints, bool flip,
                             bool bIsSparse, int imgidx) {
  std::vector<std::vector<IgnoredHyp>> best;
  computeRootPosteriorRot(part_app, log_part_detections, root_part_posterior, rootpart_idx, joints, flip, bIsSparse,
                          imgidx, best, false);
}

// computeRotJointMarginal: child [R][H][W] -> parent [R][H][W] (host grids).
template <class ExpParamT, class Grid3, class VecT, class MatT>
void computeRotJointMarginal(const ExpParamT &exp_param, Grid3 &log_prob_child, Grid3 &log_prob_parent,
                             const VecT &offset_c_10, const VecT &offset_p_01, const MatT &C, double rot_mean,
                             double rot_sigma, double scale, bool bIsSparse) {
  const int R = (int)log_prob_child.shape()[0], H = (int)log_prob_child.shape()[1],
            W = (int)log_prob_child.shape()[2];
  const size_t n = (size_t)R * H * W;
  // a one-joint model carries the exp_param; its tree is irrelevant to ps_message
  int det[2] = {1, 1}, up[2] = {0, 0}, root[2] = {1, 0};
  ps_joint j{};
  j.child_idx = 1;
  j.parent_idx = 0;
  j.C[0] = j.C[3] = 1.0;
  ps_model_desc d{2, det, up, root, 1, &j, to_exp_param(exp_param), bIsSparse ? 1 : 0};
  ps_model *m = nullptr;
  check(ps_model_create(&d, &m), "ps_model_create");
  DeviceBuffer c(n * sizeof(float)), p(n * sizeof(float));
  double oc[2] = {offset_c_10(0), offset_c_10(1)}, op[2] = {offset_p_01(0), offset_p_01(1)};
  double cc[4] = {C(0, 0), C(0, 1), C(1, 0), C(1, 1)};
  int st = PS_OK;
  hipError_t e = hipMemcpy(c.p, log_prob_child.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) st = ps_message(m, c.f(), p.f(), H, W, oc, op, cc, rot_mean, rot_sigma, scale, bIsSparse ? 1 : 0, nullptr);
  if (e == hipSuccess && st == PS_OK)
    e = hipMemcpy(log_prob_parent.data(), p.p, n * sizeof(float), hipMemcpyDeviceToHost);
  ps_model_destroy(m);
  check_hip(e, "hipMemcpy");
  check(st, "ps_message");
}

// ------------------------------------------------------------ extra unaries
// The additions findObjectImageRotJoints makes to the log unaries before
// computeRootPosteriorRot (objectdetect_findrot.cpp:896-948), with the
// reference's signatures, on host grids (each call uploads the grids it
// changes, adds on the GPU and downloads them).  On device unaries use
// ps_add_torso_pos_prior / ps_add_dpm_score / ps_add_extra_unary directly.

namespace detail {
// f(device grid of part p, scale s) for every (p, s) in `parts`
template <class Grid3, class F>
void on_device(std::vector<std::vector<Grid3>> &grids, const std::vector<int> &parts, F f) {
  for (int p : parts)
    for (auto &g : grids[p]) {
      const size_t n = g.num_elements();
      DeviceBuffer d(n * sizeof(float));
      check_hip(hipMemcpy(d.p, g.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
      f(d.f(), (int)g.shape()[0], (int)g.shape()[1], (int)g.shape()[2], p);
      check_hip(hipMemcpy(g.data(), d.p, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    }
}
template <class Grid2>
size_t slices_bytes(const std::vector<Grid2> &g) {
  if (g.empty()) throw std::invalid_argument("no DPM score grids");
  return g.size() * g[0].num_elements() * sizeof(float);
}
template <class Grid2>
void upload_slices(const std::vector<Grid2> &g, const DeviceBuffer &d) {
  const size_t hw = g[0].num_elements();
  for (size_t i = 0; i < g.size(); ++i)
    check_hip(hipMemcpy(d.f() + i * hw, g[i].data(), hw * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
}
}  // namespace detail

// setTorsoPosPrior (objectdetect_icps.cpp:137-191); MatT: pos_prior_params(0, j)
template <class PartAppT, class Grid3, class MatT>
void setTorsoPosPrior(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                      const MatT &pos_prior_params, int rootpart_idx) {
  const double params[4] = {pos_prior_params(0, 0), pos_prior_params(0, 1), pos_prior_params(0, 2),
                            pos_prior_params(0, 3)};
  const float weight = part_app.m_exp_param.torso_pos_prior_weight();
  detail::on_device(log_part_detections, {rootpart_idx}, [&](float *g, int R, int H, int W, int) {
    check(ps_add_torso_pos_prior(g, R, H, W, params, weight, nullptr), "ps_add_torso_pos_prior");
  });
}

// addDPMScore (objectdetect_icps.cpp:490-524): log grids, one or one per rotation
template <class PartAppT, class Grid3, class Grid2>
void addDPMScore(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                 std::vector<Grid2> log_dpmPriorGrid, int pidx, float dpm_weight) {
  (void)part_app;
  DeviceBuffer d(detail::slices_bytes(log_dpmPriorGrid));
  detail::upload_slices(log_dpmPriorGrid, d);
  const int n = (int)log_dpmPriorGrid.size();
  detail::on_device(log_part_detections, {pidx}, [&](float *g, int R, int H, int W, int) {
    check(ps_add_dpm_score(g, 1, R, H, W, d.f(), n, dpm_weight, PS_DPM_LOG, nullptr), "ps_add_dpm_score");
  });
}

// addLoadDPMScore (objectdetect_icps.cpp:445-488) with the scoregrid slices
// already loaded (loadDPMScoreGrid, :550-575): raw scores for every detected
// part, or for pidx_only when it names one
template <class PartAppT, class Grid3, class Grid2>
void addDPMScoreRaw(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                    const std::vector<Grid2> &dpmScoreGrid, float dpm_weight, int pidx_only = -1) {
  const int P = part_app.m_part_conf.part_size();
  std::vector<int> parts;
  for (int p = 0; p < P; ++p)
    if (!(pidx_only > -1 && pidx_only < P && pidx_only != p) && part_app.m_part_conf.part(p).is_detect())
      parts.push_back(p);
  DeviceBuffer d(detail::slices_bytes(dpmScoreGrid));
  detail::upload_slices(dpmScoreGrid, d);
  const int n = (int)dpmScoreGrid.size();
  detail::on_device(log_part_detections, parts, [&](float *g, int R, int H, int W, int) {
    check(ps_add_dpm_score(g, 1, R, H, W, d.f(), n, dpm_weight, PS_DPM_RAW, nullptr), "ps_add_dpm_score");
  });
}

// addExtraUnary (objectdetect_icps.cpp:526-548)
template <class PartAppT, class Grid3>
void addExtraUnary(const PartAppT &part_app, std::vector<std::vector<Grid3>> &log_part_detections,
                   const std::vector<std::vector<Grid3>> &log_extra_unary_scores, float weight) {
  const int P = part_app.m_part_conf.part_size();
  std::vector<int> parts(P);
  for (int p = 0; p < P; ++p) parts[p] = p;
  std::vector<size_t> next(P, 0);
  detail::on_device(log_part_detections, parts, [&](float *g, int R, int H, int W, int p) {
    const auto &e = log_extra_unary_scores[p][next[p]++];
    const size_t n = (size_t)R * H * W;
    DeviceBuffer d(n * sizeof(float));
    check_hip(hipMemcpy(d.p, e.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    check(ps_add_extra_unary(g, d.f(), (long long)n, weight, nullptr), "ps_add_extra_unary");
  });
}

}  // namespace partapp_ps
