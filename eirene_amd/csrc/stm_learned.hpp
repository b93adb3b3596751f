// stm_learned.hpp — learned leaf routing: model layout and the device-side predictor.
//
// The tree emits a leaf directory: the low fences of all leaves in key order (fences[r],
// with fences[L] = +inf) and their node ids (lids[r]).  A linear root maps a key to one of S
// segments (the line through the first and last fence, in segment units); a per-segment MLP
// 1 -> 16 -> 1 (ReLU, fp32) predicts the rank inside the segment.  The root is kept across
// periodic retrains, so each segment keeps its key range and its model a warm start.  The
// lookup verifies the window [p-2, p+2] exactly against the directory fences and the
// leaf's current high fence, and falls back to a root-to-leaf descent on a miss.
#pragma once
#include "stm_common.hpp"

namespace stm {

constexpr int MLP_HIDDEN = 16;
constexpr int MLP_PARAMS = 3 * MLP_HIDDEN + 1;  // w1[16], b1[16], w2[16], b2
constexpr int ROUTE_WINDOW = 2;                 // predicted rank +/- 2
constexpr int ROUTE_SEGMENTS = 4096;

// One segment: 256 bytes, read by the lookup with one coalesced load per lane group.
struct alignas(256) SegModel {
  uint64_t klo;       // smallest training key of the segment
  float inv_w;        // 1 / (khi - klo) in key units (x = (key - klo) * inv_w)
  float n;            // ranks covered by the segment (output scale)
  uint32_t r0;        // first rank of the segment
  uint32_t cnt;       // training pairs (ranks) in the segment
  float p[MLP_PARAMS];  // w1[0..15], b1[16..31], w2[32..47], b2[48]
  float pad[64 - 6 - MLP_PARAMS];
};
static_assert(sizeof(SegModel) == 256, "segment record is 256 bytes");

// Device-resident header of the router.
struct RouteHeader {
  uint32_t L;   // leaves in the directory
  uint32_t S;   // segments
  uint32_t L_next;  // leaves of the directory being rewritten
  uint32_t pad;
  double a, b;  // root line: segment ~ a * key + b
};

struct RouteView {
  const uint64_t* fences;  // [L + 1], fences[L] = KEY_EMPTY
  const uint32_t* lids;    // [L]
  const SegModel* segs;    // [S]
  const DirEntry* dir;     // [L + 1], dir[L].fence = KEY_EMPTY
  uint32_t L;
  uint32_t S;
  double a, b;             // root: segment ~ a * key + b
};

// MLP forward on a normalized input: returns the normalized output.
__host__ __device__ inline float mlp_forward(const float* p, float x) {
  float y = p[3 * MLP_HIDDEN];
  for (int j = 0; j < MLP_HIDDEN; ++j) {
    const float h = fmaxf(p[j] * x + p[MLP_HIDDEN + j], 0.0f);
    y += p[2 * MLP_HIDDEN + j] * h;
  }
  return y;
}

__host__ __device__ inline uint32_t root_segment(const RouteView& r, uint64_t key) {
  double y = r.a * (double)key + r.b;
  if (!(y > 0.0)) y = 0.0;  // also maps NaN to 0
  if (y >= (double)r.S) return r.S - 1;
  const uint32_t s = (uint32_t)y;
  return s >= r.S ? r.S - 1 : s;
}

STM_DEV RouteView route_view(const RouteHeader* rh, const uint64_t* fences, const uint32_t* lids,
                             const SegModel* segs, const DirEntry* dir = nullptr) {
  RouteView v;
  v.fences = fences;
  v.lids = lids;
  v.segs = segs;
  v.dir = dir;
  v.L = rh->L;
  v.S = rh->S;
  v.a = rh->a;
  v.b = rh->b;
  return v;
}

// Device buffers of the router (owned by the tree).
struct RouteState {
  RouteHeader* hdr = nullptr;
  uint64_t* fences = nullptr;  // capacity + 1
  uint32_t* lids = nullptr;    // capacity
  SegModel* segs = nullptr;    // ROUTE_SEGMENTS
  DirEntry* dir = nullptr;     // capacity + 1: packed directory read by the lookup kernel
  uint32_t* rank_of = nullptr; // capacity: node id -> directory rank (NODE_NONE if none)
  // directory build scratch
  uint64_t *lows = nullptr, *lows_tmp = nullptr, *sorted_lows = nullptr;
  uint32_t *order = nullptr, *order_tmp = nullptr, *hist = nullptr;
  DirEntry* dir_alt = nullptr;  // capacity + 1: the directory being rewritten (warm retrain)
  uint32_t builds = 0;          // directory builds so far (varies the training subsample)
};

// Builds the directory and the root; trains the segment MLPs for `steps` Adam steps when
// train is set.  warm keeps the segments' current parameters as the starting point (a
// retrain after a few batches); otherwise they restart from the identity-like initial ones.
hipError_t route_build(const TreeView& t, RouteState& rs, uint32_t n_nodes, uint32_t steps, float lr,
                       bool train, bool warm, hipStream_t s);
hipError_t mlp_grad(const float* d_x, const float* d_t, uint32_t n, const float* d_params,
                    float* d_grad, float* d_loss, hipStream_t s);
// d_hm[0] += ranks whose own fence is predicted within the window, d_hm[1] += the others.
hipError_t route_evaluate(const RouteState& rs, uint64_t* d_hm, uint32_t n_max, hipStream_t s);

}  // namespace stm
