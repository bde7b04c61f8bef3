// Internal structures shared by the host planner (ps_plan.cpp, ps_api.hip)
// and the gfx950 kernels (ps_kernels.hip).  Plain structs only: they are
// copied to device memory as tables.
#pragma once

#include <stdint.h>

namespace ps {

constexpr float kLogZero = -1e6f;  // LOG_ZERO, boost_math.h:23
constexpr float kLog2eF = 1.4426950408889634f;  // log2(e)
constexpr int kMaxRot = 64;        // rotation masks are 64-bit words
constexpr int kMaxRootChildren = 16;
constexpr int kMaxTaps = 1000;     // multi_array_filter.hpp:243-244
constexpr int kNoSourceRow = -(1 << 30);  // row offset of an output slice with no source slice
constexpr int kMfmaMaxKK = 8;      // BF16 Toeplitz passes: half widths up to 47 (2k + 32 <= 8 * 16)

// Which base pointer an offset is relative to.
enum BufId : int { kBufUnary = 0, kBufMarginal = 1, kBufScratch = 2, kBufRoot = 3 };

// A grid [R][H][pitch] of floats: the caller's unaries and marginals are
// dense (pitch W), and so is log_from_root (held in a marginal grid until the
// downward pass writes the marginal); the column passes' own outputs
// (message slots, upward beliefs, downward sources) start each row on a
// 128-byte boundary (grid_pitch(W)), and the matrix-core column passes write
// them over the whole pitch, so that every 128-byte line is written by one
// wave in one piece: lines written in parts by different waves cost the
// column passes ~5 % more time than the same bytes in whole lines.
struct GridRef {
  int buf;    // BufId, or -1 for "none"
  int pitch;  // row pitch (floats); a slice holds H rows
  long long off;  // element offset into that buffer
};

// The base pointers of one execution, passed by value to every kernel.
struct Bases {
  const float *unary;
  float *marginal;
  float *scratch;
  float *root;
  unsigned *max_keys;            // order-preserving keys of grid maxima
  unsigned long long *map_keys;  // (score, first index) argmax keys
  unsigned long long *top2_keys; // per argmax key: (best value key << 32) | second best value key, or null
  const unsigned *active;        // per instance: 0 = skip it in this execution; null = run every instance
  const unsigned char *keep;     // pruned execution: keep flag per 64 x 64 tile of every (part, rotation)
                                 // plane (MsgJob::keep_off); null = every tile
};

// Upright / border masking of one part's unaries
// (objectdetect_findrot.cpp:509-551), applied on read.
struct UnaryMask {
  unsigned long long rot_bits;  // bit r set: rotation r reads LOG_ZERO
  int strip;                    // columns [0,strip) and [W-strip,W) read LOG_ZERO
  int pad_;
};

enum SrcKind : int { kSrcGrid = 0, kSrcUnary = 1, kSrcZero = 2 };
enum DstKind : int { kDstUpNode = 0, kDstUpRoot = 1, kDstDown = 2 };
enum RotMode : int { kRotFilter = 0, kRotCopy = 1, kRotZero = 2 };

// One message (instance, scale, joint, direction) of the sweep.  A message is
// computeRotJointMarginal (objectdetect_findrot.cpp:292-456) followed by the
// accumulation its caller performs on the result.
struct MsgJob {
  int H, W;
  int inst;      // instance of the plan
  int split;     // 1: a warp's offset changes once along an axis (breaks and steps in the shift pool)
  int delta;     // rotation shift: output slice = source slice + delta
  int rot_mode;  // RotMode
  int rot_len;   // taps of the circular rotation filter
  int kx, ky;    // half widths of the x and y filters
  int taps_r;    // offsets into the tap pool (floats); x/y tap runs are padded
  int taps_x;    //   with zeros to a multiple of 16
  int taps_y;
  int shifts;    // offset into the shift pool: fwd_ox[R] fwd_oy[R] inv_ox[R] inv_oy[R] (by source
                 // slice), then fwd_ox[R] fwd_oy[R] by output slice, then those packed in 16 bits each,
                 // then the clamped source slice by output slice; split: from 8 R the forward x break,
                 // x step, y break, y step by output slice, then the inverse ones by slice
  int src_kind;  // SrcKind
  int src_key;   // index of the source's max key
  GridRef src;   // [R][H][W]
  UnaryMask src_mask;
  GridRef tmp;   // row-pass output [R][H][tpitch]
  // destination
  int dst_kind;       // DstKind
  int dst_key;        // kDstUpNode: key of the parent belief (next source)
  GridRef dst;        // UpNode: parent belief; UpRoot: message slot; Down: child marginal
  GridRef dst_unary;  // unary of the part written (UpNode: parent, Down: child)
  UnaryMask dst_mask;
  int dst_add_unary;  // UpNode: parent is_detect
  int dst_base;       // Down: 0 = the stored upward belief (dst_bel), 1 = masked unary, 2 = zero
  GridRef dst2;       // Down with a grandchild: unary + message (next source), else buf -1
  GridRef dst_bel;    // Down with a grandchild: the stored upward belief (pitch of dst2), else buf -1
  int dst2_key;
  int map_slot;       // Down: argmax key of the child marginal
  // BF16 Toeplitz variant of the x / y passes: k-steps of 16 and offsets (in
  // 16-byte units) of the tap fragment tables (hi block, then lo block)
  int kk_x, kk_y;
  int frag_x, frag_y;
  int tpitch;  // row pitch (floats) of the row-pass output grid, a multiple of 32
  int frag_r;  // rotation filter as the A operand of the row pass's second GEMM
  int t16_shift;  // BF16 passes with an fp16 row-pass output: stored as value * 2^t16_shift
  int frag_y16;   // the y taps as fp16 hi / lo tap copies (column pass on an fp16 row-pass output)
  int frag_r16;   // the rotation filter as fp16 hi / lo A fragments (the row pass with an fp16 output)
  int x16_shift;  // that row pass exponentiates to value * 2^x16_shift (its x filter's largest output near 2^14)
  // Off-diagonal covariance (gaussFilter2dOffset, multi_array_filter.hpp:
  // 334-369): the row pass writes the rotation-filtered grid only (kx = 0, a
  // unit x tap); each slice is then rotated onto a th x tw grid in the
  // eigenbasis of scale^2 C, filtered there with the eigenvalue taps and
  // sampled back bilinearly.
  int od;          // 1: this message takes the eigenbasis path
  int od_direct;   // forward rotation: 1 nearest scatter (TM_DIRECT), 0 bilinear gather
  int th, tw;      // eigenbasis grid
  int kxe, kye;    // half widths of the eigenbasis filters
  int taps_xe;     // offsets into the tap pool
  int taps_ye;
  GridRef od_buf;  // floats: A[R][th][tw] (scatter keys: 2 floats per cell), then X[R][th][tw]
  // pruned execution: the destination part's keep flags, [R][keep_ty][keep_tx]
  // tiles of 64 x 64 starting at keep_off (bytes) in Bases::keep
  long long keep_off;
  int keep_ty, keep_tx;
  double t_fwd[6];   // rows 0-1; direct: grid -> eigenbasis, bilinear: eigenbasis -> grid
  double t_back[6];  // rows 0-1; grid pixel -> eigenbasis sample position
};

// The caller's accumulation of a message's column pass
// (objectdetect_findrot.cpp:637-654 upward, :201 and :220-222 downward), in
// increasing order of the grids it touches
enum EpiKind : int {
  kEpiUpRoot = 0,    // message slot of the root: msg
  kEpiUpNode = 1,    // chain-node belief: msg
  kEpiUpNodeU = 2,   // chain-node belief: msg + masked unary
  kEpiDownLeaf = 3,  // marginal of a leaf: (masked unary or 0) + msg
  kEpiDownMid = 4,   // marginal: stored belief + msg; next downward source: masked unary + msg
};
constexpr int epi_kind(const MsgJob &J) {
  return J.dst_kind == kDstUpRoot ? kEpiUpRoot
         : J.dst_kind == kDstUpNode ? (J.dst_add_unary ? kEpiUpNodeU : kEpiUpNode)
         : J.dst_base != 0 ? kEpiDownLeaf
                           : kEpiDownMid;
}

// The y taps of the column pass on the fp16 MFMA (k-steps KK) as 8 copies of
// the zero-padded taps g[p] = f[p - 32], copy j holding g[j], g[j + 1], ...:
// the 8 taps of any lane's A fragment are then one aligned 16-byte read.
// Elements per copy: at least 16 KK + 32, padded to a byte length of 64
// modulo 256 so that the copies start in different LDS banks.
constexpr int f16_tap_copy_len(int KK) {
  return (256 * ((2 * (16 * KK + 32) - 64 + 255) / 256) + 64) / 2;
}

// floats of the eigenbasis scratch of one message
inline long long od_scratch_floats(int R, int th, int tw) { return 3ll * R * th * tw; }
// row pitch of the row-pass output and of the plan's own grids: rows start
// on 128-byte boundaries
inline int t_pitch(int W) { return (W + 31) & ~31; }
inline int grid_pitch(int W) { return t_pitch(W); }

// Root of one (instance, scale): sums the upward messages into the root
// belief, forms the "all but one child" root messages
// (objectdetect_findrot.cpp:637-649, 157-170) and the root posterior (694-726).
struct RootJob {
  int H, W;
  int K;  // root children, in joint order
  int add_unary;
  int inst;
  int pad0_;
  GridRef msg[kMaxRootChildren];  // upward messages (the plan's grids, one row pitch for all)
  GridRef fr[kMaxRootChildren];   // log_from_root of each child (dense: the child's marginal grid)
  int fr_key[kMaxRootChildren];
  GridRef unary;
  UnaryMask mask;
  GridRef marginal;  // root marginal (= belief after the upward pass)
  int map_slot;
  int pad_;
  GridRef root_post;  // [H][W] of this scale, buf -1 when not wanted
  unsigned long long valid_rot;  // rotations summed by the root posterior
};

// Max of a leaf source grid.
struct MaxJob {
  int H, W;
  int src_kind;
  int key;
  int inst;
  int pad0_;
  GridRef src;
  UnaryMask mask;
};

// Decoding of one argmax key into a ps_map record.
struct MapJob {
  int H, W;
  int slot;
  int out;  // index into the ps_map array
  int inst;
  int pad0_;
};

}  // namespace ps
