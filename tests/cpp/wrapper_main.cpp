// Drives include/partapp_ps.hpp with stand-ins of partapp's types (PartApp
// with protobuf-style accessors, uBLAS-style Joint, a multi_array-like grid,
// PartHyp), the way partapp's findObjectImageRotJoints calls
// computeRootPosteriorRot.  Reads the log unaries [P][S][R][H][W] and joints
// from a binary file, writes the root posterior, the log marginals, the
// best_part_hyp records, the masked unaries, best_conf and the root
// HypothesisList.
//   wrapper_main <in.bin> <out.bin>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "partapp_ps.hpp"

struct ExpParam {
  float min_rot, max_rot, min_scale, max_scale, strip;
  unsigned R, S;
  float min_part_rotation() const { return min_rot; }
  float max_part_rotation() const { return max_rot; }
  unsigned num_rotation_steps() const { return R; }
  float min_object_scale() const { return min_scale; }
  float max_object_scale() const { return max_scale; }
  unsigned num_scale_steps() const { return S; }
  float strip_border_detections() const { return strip; }
  float roi_save_num_samples() const { return 1000.0f; }
};
struct PartDef {
  bool det, up, root;
  bool is_detect() const { return det; }
  bool is_upright() const { return up; }
  bool is_root() const { return root; }
};
struct PartConfig {
  std::vector<PartDef> parts;
  int part_size() const { return (int)parts.size(); }
  const PartDef &part(int p) const { return parts[p]; }
};
struct WindowParam {
  double bbox_offset_x() const { return -12.7; }
  double bbox_offset_y() const { return 5.3; }
};
struct PartApp {
  ExpParam m_exp_param;
  PartConfig m_part_conf;
  WindowParam m_window_param;
};
// protobuf-style HypothesisList (HypothesisList.proto)
struct ObjectHypothesis {
  float x_ = 0, y_ = 0, scale_ = 0, score_ = 0;
  bool flip_ = false;
  float x() const { return x_; }
  float y() const { return y_; }
  void set_x(float v) { x_ = v; }
  void set_y(float v) { y_ = v; }
  void set_scale(float v) { scale_ = v; }
  void set_score(float v) { score_ = v; }
  void set_flip(bool v) { flip_ = v; }
};
struct HypothesisList {
  std::vector<ObjectHypothesis> h;
  ObjectHypothesis *add_hyp() {
    h.emplace_back();
    return &h.back();
  }
  int hyp_size() const { return (int)h.size(); }
  const ObjectHypothesis &hyp(int i) const { return h[i]; }
  ObjectHypothesis *mutable_hyp(int i) { return &h[i]; }
};
struct Vec2 {
  double v[2];
  double operator()(int i) const { return v[i]; }
};
struct Mat2 {
  double m[4];
  double operator()(int i, int j) const { return m[2 * i + j]; }
};
struct Joint {
  int child_idx, parent_idx;
  Vec2 offset_c, offset_p;
  Mat2 C;
  double rot_mean, rot_sigma;
};
struct Grid3 {
  std::array<size_t, 3> ext{0, 0, 0};
  std::vector<float> d;
  float *data() { return d.data(); }
  const float *data() const { return d.data(); }
  const size_t *shape() const { return ext.data(); }
  void resize(const std::array<size_t, 3> &e) {
    ext = e;
    d.resize(e[0] * e[1] * e[2]);
  }
};
struct PartHyp {
  int scaleidx, rotidx, x, y;
  float score;
  double scale, rot;
  PartHyp(const ExpParam &e, int s, int r, int x_, int y_, float sc) : scaleidx(s), rotidx(r), x(x_), y(y_), score(sc) {
    scale = e.min_scale == e.max_scale ? (double)e.min_scale
                                       : (double)e.min_scale + ((double)e.max_scale - e.min_scale) / e.S * (0.5 + s);
    rot = e.min_rot == e.max_rot ? (double)e.min_rot
                                 : (double)e.min_rot + ((double)e.max_rot - e.min_rot) / e.R * (0.5 + r);
  }
  std::array<float, 7> toVect() const {
    return {(float)scaleidx, (float)scale, (float)rotidx, (float)rot, (float)x, (float)y, score};
  }
};

template <class T>
static void rd(FILE *f, T *p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[6];  // P S R H W root
  rd(f, hdr, 6);
  const int P = hdr[0], S = hdr[1], R = hdr[2], H = hdr[3], W = hdr[4], root = hdr[5];
  PartApp app;
  float ep[5];
  rd(f, ep, 5);
  app.m_exp_param = ExpParam{ep[0], ep[1], ep[2], ep[3], ep[4], (unsigned)R, (unsigned)S};
  std::vector<int> flags(3 * P);
  rd(f, flags.data(), flags.size());
  for (int p = 0; p < P; ++p) app.m_part_conf.parts.push_back(PartDef{flags[p] != 0, flags[P + p] != 0, flags[2 * P + p] != 0});
  std::vector<Joint> joints(P - 1);
  for (auto &j : joints) {
    int ids[2];
    double v[10];
    rd(f, ids, 2);
    rd(f, v, 10);
    j = Joint{ids[0], ids[1], Vec2{{v[0], v[1]}}, Vec2{{v[2], v[3]}}, Mat2{{v[4], v[5], v[6], v[7]}}, v[8], v[9]};
  }
  std::vector<std::vector<Grid3>> det(P, std::vector<Grid3>(S));
  for (int p = 0; p < P; ++p)
    for (int s = 0; s < S; ++s) {
      det[p][s].resize({(size_t)R, (size_t)H, (size_t)W});
      rd(f, det[p][s].data(), (size_t)R * H * W);
    }
  fclose(f);

  Grid3 root_post;
  std::vector<std::vector<PartHyp>> best;
  std::vector<std::vector<Grid3>> marg;
  HypothesisList hyps;
  std::vector<float> best_conf;
  try {
    partapp_ps::computeRootPosteriorRot(app, det, root_post, root, joints, false, true, 0, best, false, &marg);
    partapp_ps::rootHypotheses(app, root_post, true, hyps);
    best_conf = partapp_ps::bestConf(best);
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(root_post.data(), sizeof(float), (size_t)S * H * W, o);
  for (int p = 0; p < P; ++p)
    for (int s = 0; s < S; ++s) fwrite(marg[p][s].data(), sizeof(float), (size_t)R * H * W, o);
  for (int p = 0; p < P; ++p) {
    int nh = (int)best[p].size();
    fwrite(&nh, sizeof(int), 1, o);
    for (const PartHyp &h : best[p]) {
      int rec[4] = {h.scaleidx, h.rotidx, h.y, h.x};
      fwrite(rec, sizeof(int), 4, o);
      fwrite(&h.score, sizeof(float), 1, o);
    }
  }
  for (int p = 0; p < P; ++p)
    for (int s = 0; s < S; ++s) fwrite(det[p][s].data(), sizeof(float), (size_t)R * H * W, o);
  fwrite(best_conf.data(), sizeof(float), best_conf.size(), o);
  const int nh = hyps.hyp_size();
  fwrite(&nh, sizeof(int), 1, o);
  for (const ObjectHypothesis &h : hyps.h) {
    float rec[5] = {h.x_, h.y_, h.scale_, h.score_, h.flip_ ? 1.0f : 0.0f};
    fwrite(rec, sizeof(float), 5, o);
  }
  fclose(o);
  printf("ok\n");
  return 0;
}
