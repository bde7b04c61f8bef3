// Seeded synthetic instances: the benchmark workload of the sweep.
//
// Random numbers: the i-th output of a SplitMix64 stream started at state s
// is mix64(s + (i + 1) * 0x9E3779B97F4A7C15), so every element can be drawn
// independently.  Uniforms are (top 53 bits + 1) * 2^-53, in (0, 1];
// normals come from Box-Muller in double, two per pair of uniforms.
//
// Person tree (10 parts, root = part 4, the torso), joints in the order of
// src/scripts/matlab/getJointsParts.m:17-25 with 0-based part indices:
//   0->1, 1->4, 2->4, 3->2, 5->4, 6->7, 7->4, 8->4, 9->8.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/partapp_ps.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_at(uint64_t state, uint64_t i) { return mix64(state + (i + 1) * kGamma); }
inline double u01(uint64_t z) { return (double)((z >> 11) + 1) * 0x1.0p-53; }

struct Stream {
  uint64_t state, i = 0;
  explicit Stream(uint64_t s) : state(s) {}
  double uniform() { return u01(stream_at(state, i++)); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int index(int n) {
    int v = (int)std::floor(uniform() * n);
    return v >= n ? n - 1 : v;
  }
};

uint64_t key(uint64_t seed, uint64_t tag, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed ^ tag) + a) + b);
}

constexpr uint64_t kTagJoints = 0x6a6f696e7473ull;  // "joints"
constexpr uint64_t kTagBumps = 0x62756d7073ull;     // "bumps"
constexpr uint64_t kTagUnary = 0x756e617279ull;     // "unary"

const int kPersonChild[9] = {0, 1, 2, 3, 5, 6, 7, 8, 9};
const int kPersonParent[9] = {1, 4, 4, 2, 4, 7, 4, 4, 8};

}  // namespace

extern "C" int ps_synth_joints(unsigned long long seed, int tree, int zero_rot_mean, ps_joint *joints,
                               int *num_joints) {
  if (tree != 0 || !joints) return PS_ERR_INVALID_ARG;
  Stream st(key(seed, kTagJoints, 0, 0));
  for (int j = 0; j < 9; ++j) {
    ps_joint &J = joints[j];
    J.child_idx = kPersonChild[j];
    J.parent_idx = kPersonParent[j];
    J.offset_c[0] = st.uniform(-40.0, 40.0);
    J.offset_c[1] = st.uniform(-40.0, 40.0);
    J.offset_p[0] = st.uniform(-40.0, 40.0);
    J.offset_p[1] = st.uniform(-40.0, 40.0);
    double sx = st.uniform(4.0, 12.0), sy = st.uniform(4.0, 12.0);
    J.C[0] = sx * sx;
    J.C[1] = 0.0;
    J.C[2] = 0.0;
    J.C[3] = sy * sy;
    J.rot_sigma = st.uniform(0.1, 0.4);
    double rm = st.uniform(-M_PI / 6, M_PI / 6);
    J.rot_mean = zero_rot_mean ? 0.0 : rm;
  }
  if (num_joints) *num_joints = 9;
  return PS_OK;
}

extern "C" int ps_synth_unary(unsigned long long seed, int P, int S, int R, int H, int W, float *unary,
                              int n_threads) {
  if (P < 1 || S < 1 || R < 1 || H < 1 || W < 1 || !unary) return PS_ERR_INVALID_ARG;
  const int kBumps = 3;
  std::vector<int> bx(P * kBumps), by(P * kBumps), br(P * kBumps);
  for (int p = 0; p < P; ++p) {
    Stream st(key(seed, kTagBumps, (uint64_t)p, 0));
    for (int k = 0; k < kBumps; ++k) {
      bx[p * kBumps + k] = st.index(W);
      by[p * kBumps + k] = st.index(H);
      br[p * kBumps + k] = st.index(R);
    }
  }
  const long long plane = (long long)H * W, n = (long long)R * plane;
  const long long rows = (long long)P * S * R * H;
  auto work = [&](long long row_begin, long long row_end) {
    for (long long row = row_begin; row < row_end; ++row) {
      const int y = (int)(row % H);
      const int r = (int)((row / H) % R);
      const int s = (int)((row / ((long long)H * R)) % S);
      const int p = (int)(row / ((long long)H * R * S));
      const uint64_t state = key(seed, kTagUnary, (uint64_t)p, (uint64_t)s);
      float *out = unary + ((long long)p * S + s) * n + (long long)r * plane + (long long)y * W;
      for (int x = 0; x < W; ++x) {
        const long long e = (long long)r * plane + (long long)y * W + x;
        const long long pair = e >> 1;
        const double u1 = u01(stream_at(state, 2 * pair));
        const double u2 = u01(stream_at(state, 2 * pair + 1));
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double th = 2.0 * M_PI * u2;
        const double z = (e & 1) ? rad * std::sin(th) : rad * std::cos(th);
        double v = -3.0 + 0.5 * z;
        for (int k = 0; k < kBumps; ++k) {
          const int q = p * kBumps + k;
          if (br[q] != r) continue;
          const double dx = x - bx[q], dy = y - by[q];
          v += 4.0 * std::exp(-(dx * dx + dy * dy) / (2.0 * 36.0));
        }
        out[x] = (float)v;
      }
    }
  };
  int nt = n_threads > 0 ? n_threads : 1;
  if (nt > 64) nt = 64;
  if (nt == 1 || rows < 64) {
    work(0, rows);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, rows * t / nt, rows * (t + 1) / nt);
    for (auto &x : th) x.join();
  }
  return PS_OK;
}
