// stm_learned.hip — leaf directory and GPU training of the learned leaf router.
//
// No reference counterpart: the reference routes every op from the root (v38) or through a
// rebuilt static index (v35, src/gpu-stm-tree-v35/bpt-gpustm-kernel.cu:17-231).  Here the
// tree emits (key, leaf rank) pairs — the low fence of every leaf in key order — and the
// router is trained on them on the device:
//   k_route_directory     leaves sorted by low fence -> fences[r], lids[r], L
//   k_route_root          root line segment ~ a*key + b through the first and last fence
//   k_route_segments      segment s = ranks whose root prediction falls in bucket s
//   k_route_train         one wavefront per segment: full-batch Adam on the MSE of a
//                         1 -> 16 -> 1 ReLU MLP (hand-written forward and backward)
// k_mlp_grad exposes one forward/backward pass for the gradient check against the CPU.
#include <cmath>

#include "stm_common.hpp"
#include "stm_internal.hpp"
#include "stm_learned.hpp"
#include "stm_tree.hpp"

namespace stm {


// Sort keys of the directory build: a leaf's low fence, +inf for every other node, so the
// leaves come first in key order after the sort (no compaction and no shared counter).
__global__ void k_route_lows(TreeView t, uint32_t n_nodes, uint64_t* __restrict__ lows,
                             uint32_t* __restrict__ rank_of) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < t.capacity) rank_of[i] = NODE_NONE;
  if (i < n_nodes) {
    const NodeMeta m = t.meta[i];
    lows[i] = meta_level(m.info) == 0 ? m.low : KEY_EMPTY;
  }
}

// Directory from the sorted (low fence, node id) pairs.  The last leaf writes L.
__global__ void k_route_directory(const uint64_t* __restrict__ sorted_lows,
                                  const uint32_t* __restrict__ order, uint32_t n,
                                  RouteHeader* __restrict__ rh, uint64_t* __restrict__ fences,
                                  uint32_t* __restrict__ lids, DirEntry* __restrict__ dir,
                                  uint32_t* __restrict__ rank_of) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint64_t low = sorted_lows[r];
  if (low == KEY_EMPTY) return;
  fences[r] = low;
  lids[r] = order[r];
  dir[r] = DirEntry{low, order[r], DIR_PAR_NONE};
  rank_of[order[r]] = r;
  if (r + 1 == n || sorted_lows[r + 1] == KEY_EMPTY) {
    rh->L = r + 1;
    rh->S = ROUTE_SEGMENTS;
    fences[r + 1] = KEY_EMPTY;
    dir[r + 1] = DirEntry{KEY_EMPTY, NODE_NONE, DIR_SPLIT | DIR_PAR_NONE};
  }
}

// Parent hints: every level-1 node writes its id into its children's directory entries.
__global__ void k_route_parents(TreeView t, uint32_t n_nodes, DirEntry* __restrict__ dir) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n_nodes) return;
  const NodeMeta m = t.meta[id];
  if (meta_level(m.info) != 1) return;
  const uint32_t cnt = min(meta_count(m.info), t.fanout);
  const uint32_t par = id < DIR_PAR_NONE ? id : DIR_PAR_NONE;
  for (uint32_t s = 0; s < cnt; ++s) {
    const uint64_t child = t.vals[(uint64_t)id * t.fanout + s];
    if (child >= t.capacity) continue;
    const uint32_t r = t.rank_of[child];
    if (r != NODE_NONE) dir[r].pflags = par;
  }
}

// ------------------------------------------------------------------ warm directory rewrite
// Since the last directory build, leaves only split: every new leaf sits in the right-sibling
// chain that starts at an entry flagged DIR_SPLIT and ends before the next entry's leaf.  The
// rewrite counts each chain, ranks the chains by an exclusive scan and writes the new entries
// in place of a sort of every node.
constexpr int DS_BLOCK = 256;
constexpr int DS_ITEMS = 16;
constexpr int DS_TILE = DS_BLOCK * DS_ITEMS;

__global__ void k_dir_count(TreeView t, const RouteHeader* __restrict__ rh,
                            const DirEntry* __restrict__ dir, uint32_t* __restrict__ counts) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t L = rh->L;
  if (r >= L) return;
  const DirEntry e = dir[r];
  uint32_t c = 1;
  if (e.pflags & DIR_SPLIT) {
    const uint32_t stop = r + 1 < L ? dir[r + 1].lid : NODE_NONE;
    uint32_t node = e.lid;
    for (uint32_t guard = 0; guard < t.capacity; ++guard) {
      const uint64_t nx = t.meta[node].next;
      if (nx == stop || nx >= t.capacity) break;
      node = (uint32_t)nx;
      ++c;
    }
  }
  counts[r] = c;
}

STM_DEV uint32_t block_exclusive_u32(uint32_t v, uint32_t* s_w /*[4]*/, uint32_t& total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, WAVE);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) s_w[wave] = x;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < DS_BLOCK / WAVE; ++w) {
    const uint32_t t = s_w[w];
    if (w < wave) before += t;
    total += t;
  }
  __syncthreads();
  return before + x - v;
}

// Exclusive scan of counts[0 .. L) into offs (three kernels: tile sums, their scan, apply).
__global__ __launch_bounds__(DS_BLOCK) void k_dir_scan_reduce(const RouteHeader* __restrict__ rh,
                                                              const uint32_t* __restrict__ counts,
                                                              uint32_t* __restrict__ tiles) {
  __shared__ uint32_t s_w[DS_BLOCK / WAVE];
  const uint32_t L = rh->L;
  const uint32_t b = blockIdx.x;
  if ((uint64_t)b * DS_TILE >= L) return;  // uniform over the block
  uint32_t acc = 0;
  for (int i = 0; i < DS_ITEMS; ++i) {
    const uint64_t r = (uint64_t)b * DS_TILE + (uint64_t)i * DS_BLOCK + threadIdx.x;
    if (r < L) acc += counts[r];
  }
  uint32_t total;
  (void)block_exclusive_u32(acc, s_w, total);
  if (threadIdx.x == 0) tiles[b] = total;
}

__global__ __launch_bounds__(DS_BLOCK) void k_dir_scan_tiles(RouteHeader* __restrict__ rh,
                                                             uint32_t* __restrict__ tiles) {
  __shared__ uint32_t s_w[DS_BLOCK / WAVE];
  const uint32_t nt = (rh->L + DS_TILE - 1) / DS_TILE;
  const uint32_t per = (nt + DS_BLOCK - 1) / DS_BLOCK;
  const uint32_t b0 = min(threadIdx.x * per, nt);
  const uint32_t b1 = min(b0 + per, nt);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += tiles[b];
  uint32_t total;
  uint32_t run = block_exclusive_u32(sum, s_w, total);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t c = tiles[b];
    tiles[b] = run;
    run += c;
  }
  if (threadIdx.x == 0) rh->L_next = total;
}

__global__ __launch_bounds__(DS_BLOCK) void k_dir_scan_apply(const RouteHeader* __restrict__ rh,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ tiles,
                                                             uint32_t* __restrict__ offs) {
  __shared__ uint32_t s_w[DS_BLOCK / WAVE];
  const uint32_t L = rh->L;
  const uint32_t b = blockIdx.x;
  if ((uint64_t)b * DS_TILE >= L) return;
  uint32_t carry = tiles[b];
  for (int i = 0; i < DS_ITEMS; ++i) {
    const uint64_t r = (uint64_t)b * DS_TILE + (uint64_t)i * DS_BLOCK + threadIdx.x;
    const uint32_t c = r < L ? counts[r] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_u32(c, s_w, total);
    if (r < L) offs[r] = carry + ex;
    carry += total;
  }
}

// New entries of chain r at offs[r] ..: the chain's first leaf keeps its fence; the others take
// their low fence from the node; all keep the old entry's parent hint.
__global__ void k_dir_rewrite(TreeView t, const RouteHeader* __restrict__ rh,
                              const DirEntry* __restrict__ dir, const uint32_t* __restrict__ counts,
                              const uint32_t* __restrict__ offs, DirEntry* __restrict__ dir_alt,
                              uint32_t* __restrict__ rank_of) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rh->L) return;
  const DirEntry e = dir[r];
  const uint32_t par = e.pflags & DIR_PAR_NONE;
  const uint32_t base = offs[r];
  const uint32_t c = counts[r];
  uint32_t node = e.lid;
  for (uint32_t j = 0; j < c; ++j) {
    const uint64_t low = j == 0 ? e.fence : t.meta[node].low;
    dir_alt[base + j] = DirEntry{low, node, par};
    rank_of[node] = base + j;
    if (j + 1 < c) node = (uint32_t)t.meta[node].next;
  }
}

__global__ void k_dir_install(RouteHeader* __restrict__ rh, const DirEntry* __restrict__ dir_alt,
                              DirEntry* __restrict__ dir, uint64_t* __restrict__ fences,
                              uint32_t* __restrict__ lids) {
  const uint32_t L = rh->L_next;
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L) {
    const DirEntry e = dir_alt[r];
    dir[r] = e;
    fences[r] = e.fence;
    lids[r] = e.lid;
  } else if (r == L) {
    dir[L] = DirEntry{KEY_EMPTY, NODE_NONE, DIR_SPLIT | DIR_PAR_NONE};
    fences[L] = KEY_EMPTY;
  }
}

__global__ void k_dir_set_len(RouteHeader* __restrict__ rh) {
  if (threadIdx.x == 0) rh->L = rh->L_next;
}

// The root: the line through the first and last fence, in segment units.  A least-squares
// line of a skewed key set would leave a large share of the keys left of its zero crossing,
// and clamping them into segment 0 would make that one segment hold most of the leaves.  Warm
// retrains keep the root (keys beyond the old last fence go to the last segment).
__global__ void k_route_root(const uint64_t* __restrict__ fences, RouteHeader* __restrict__ rh) {
  if (threadIdx.x != 0) return;
  const uint32_t L = rh->L;
  double a = 0, b = 0;
  if (L >= 2) {
    const double k0 = (double)fences[0], k1 = (double)fences[L - 1];
    if (k1 > k0) {
      a = (double)ROUTE_SEGMENTS / (k1 - k0);
      b = -a * k0;
    }
  }
  rh->a = a;  // segments per key unit
  rh->b = b;
}

// First rank whose root bucket is >= s (the root is monotone in the key).
STM_DEV uint32_t first_rank_of_segment(const RouteView& v, uint32_t s) {
  uint32_t lo = 0, hi = v.L;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (root_segment(v, v.fences[mid]) >= s) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

STM_DEV void mlp_init(float* p) {
  // Identity-like start: hinges at j/16, the first unit passes x through.
  for (int j = 0; j < MLP_HIDDEN; ++j) {
    p[j] = 1.0f;
    p[MLP_HIDDEN + j] = -(float)j / (float)MLP_HIDDEN;
    p[2 * MLP_HIDDEN + j] = j == 0 ? 1.0f : 0.0f;
  }
  p[3 * MLP_HIDDEN] = 0.0f;
}

__global__ void k_route_segments(RouteHeader* __restrict__ rh, const uint64_t* __restrict__ fences,
                                 const uint32_t* __restrict__ lids, SegModel* __restrict__ segs,
                                 uint32_t init_params) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const RouteView v = route_view(rh, fences, lids, segs);
  if (s >= v.S) return;
  const uint32_t r0 = first_rank_of_segment(v, s);
  const uint32_t r1 = (s + 1 < v.S) ? first_rank_of_segment(v, s + 1) : v.L;
  SegModel& m = segs[s];
  m.r0 = r0;
  m.cnt = r1 > r0 ? r1 - r0 : 0;
  m.n = (float)m.cnt;
  if (m.cnt > 0) {
    m.klo = fences[r0];
    const uint64_t khi = fences[r1 - 1];
    m.inv_w = khi > m.klo ? (float)(1.0 / ((double)khi - (double)m.klo)) : 0.0f;
  } else {
    m.klo = r0 < v.L ? fences[r0] : 0;
    m.inv_w = 0.0f;
  }
  if (init_params) mlp_init(m.p);
}

// Normalized training pair i of a segment: x in [0, 1], target (rank - r0) / cnt.
STM_DEV float seg_x(const SegModel& m, uint64_t key) {
  return (float)(((double)key - (double)m.klo) * (double)m.inv_w);
}

// Adds the gradient of 0.5 * (y - t)^2 * scale at one pair into g (per lane).
STM_DEV float mlp_accumulate(const float* p, float x, float t, float scale, float* g) {
  float z[MLP_HIDDEN];
  float y = p[3 * MLP_HIDDEN];
#pragma unroll
  for (int j = 0; j < MLP_HIDDEN; ++j) {
    z[j] = p[j] * x + p[MLP_HIDDEN + j];
    y += p[2 * MLP_HIDDEN + j] * fmaxf(z[j], 0.0f);
  }
  const float e = y - t;
  const float dy = e * scale;
  g[3 * MLP_HIDDEN] += dy;
#pragma unroll
  for (int j = 0; j < MLP_HIDDEN; ++j) {
    const float h = fmaxf(z[j], 0.0f);
    g[2 * MLP_HIDDEN + j] += dy * h;
    const float dz = z[j] > 0.0f ? dy * p[2 * MLP_HIDDEN + j] : 0.0f;
    g[j] += dz * x;
    g[MLP_HIDDEN + j] += dz;
  }
  return 0.5f * e * e * scale;
}

// Gradient reduction without LDS: a recursive-halving reduce-scatter over the wavefront.  At
// each of six stages (lane distance 32 .. 1) a lane keeps one half of its values, sends the
// other half to its partner and adds what the partner sent; after the last stage lane L holds
// the wavefront's sum of parameter param_of_lane(L) (lanes that map past MLP_PARAMS hold
// padding).  52 shuffles instead of 64 x 49 LDS words per step.
constexpr int RS_C0 = MLP_PARAMS;        // 49
constexpr int RS_H0 = (RS_C0 + 1) / 2;   // 25 values after the distance-32 stage
constexpr int RS_H1 = (RS_H0 + 1) / 2;   // 13
constexpr int RS_H2 = (RS_H1 + 1) / 2;   // 7
constexpr int RS_H3 = (RS_H2 + 1) / 2;   // 4
constexpr int RS_H4 = (RS_H3 + 1) / 2;   // 2
constexpr int RS_H5 = (RS_H4 + 1) / 2;   // 1

__host__ __device__ constexpr int param_of_lane(int lane) {
  return ((lane & 32) ? RS_H0 : 0) + ((lane & 16) ? RS_H1 : 0) + ((lane & 8) ? RS_H2 : 0) +
         ((lane & 4) ? RS_H3 : 0) + ((lane & 2) ? RS_H4 : 0) + ((lane & 1) ? RS_H5 : 0);
}
__host__ __device__ constexpr int lane_of_param(int q) {
  int lane = 0, c = RS_C0;
  for (int dist = 32; dist >= 1; dist >>= 1) {
    const int h = (c + 1) / 2;
    if (q >= h) {
      lane |= dist;
      q -= h;
    }
    c = h;
  }
  return lane;
}
static_assert(param_of_lane(lane_of_param(0)) == 0 && param_of_lane(lane_of_param(48)) == 48 &&
                  param_of_lane(lane_of_param(24)) == 24 && param_of_lane(lane_of_param(25)) == 25,
              "reduce-scatter lane map");

template <int C, int DIST>
STM_DEV void rs_stage(float* a, int lane) {
  constexpr int H = (C + 1) / 2;
  const bool up = (lane & DIST) != 0;
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const float lo = a[q];
    const float hi = q + H < C ? a[q + H] : 0.0f;
    const float send = up ? lo : hi;
    const float keep = up ? hi : lo;
    a[q] = keep + __shfl_xor(send, DIST, WAVE);
  }
}

// g[0..MLP_PARAMS) per lane -> the wavefront sum of parameter param_of_lane(lane).
STM_DEV float wave_reduce_scatter(float* g) {
  const int lane = lane_id();
  rs_stage<RS_C0, 32>(g, lane);
  rs_stage<RS_H0, 16>(g, lane);
  rs_stage<RS_H1, 8>(g, lane);
  rs_stage<RS_H2, 4>(g, lane);
  rs_stage<RS_H3, 2>(g, lane);
  rs_stage<RS_H4, 1>(g, lane);
  return g[0];
}

// Every parameter from the lane that holds it, as wave-uniform values (scalar registers).
STM_DEV void broadcast_params(float pv, float* p) {
#pragma unroll
  for (int q = 0; q < MLP_PARAMS; ++q)
    p[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pv), lane_of_param(q)));
}

constexpr int TRAIN_WAVES = 4;
constexpr int TRAIN_CACHE = 8;  // pairs per lane kept in registers
constexpr uint32_t TRAIN_MAX_PAIRS = WAVE * TRAIN_CACHE;  // 512 pairs per segment and step

// One wavefront per segment: full-batch Adam for `steps` steps.  Lane L owns parameter
// param_of_lane(L): its Adam moments and its value; the forward and backward passes read the
// parameters as wave-uniform values.
__global__ __launch_bounds__(TRAIN_WAVES * WAVE) void k_route_train(
    const RouteHeader* __restrict__ rh, const uint64_t* __restrict__ fences,
    SegModel* __restrict__ segs, uint32_t steps, float lr, uint32_t salt) {
  const int wave = threadIdx.x / WAVE;
  const int lane = lane_id();
  const uint32_t s = blockIdx.x * TRAIN_WAVES + wave;
  if (s >= rh->S) return;  // whole wavefront leaves together
  SegModel& m = segs[s];
  const uint32_t cnt = m.cnt;
  if (cnt < 2 || m.inv_w == 0.0f) return;
  const int own = param_of_lane(lane);
  const bool owner = own < MLP_PARAMS;
  float pv = owner ? m.p[own] : 0.0f;
  float am = 0.0f, av = 0.0f;  // Adam moments of the owned parameter
  const float inv_n = 1.0f / (float)cnt;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  float b1t = 1.0f, b2t = 1.0f;
  // The training pairs sit in registers for all steps.  A segment of at most TRAIN_MAX_PAIRS
  // leaves trains on all of them; a larger one (the linear root leaves dense key ranges with
  // several times the mean) on every span-th pair, from an offset that differs between builds
  // (salt), so successive retrains see different pairs and the largest segments cost no more
  // per step than a full one.
  const uint32_t span = (cnt + TRAIN_MAX_PAIRS - 1) / TRAIN_MAX_PAIRS;
  const uint32_t shift = span > 1 ? (salt * 2654435761u) % span : 0u;
  const uint32_t used = (cnt - shift + span - 1) / span;  // pairs shift, shift + span, ...
  const float scale = 1.0f / (float)used;
  float xc[TRAIN_CACHE], tc[TRAIN_CACHE];
#pragma unroll
  for (int c = 0; c < TRAIN_CACHE; ++c) {
    const uint32_t j = (uint32_t)lane + (uint32_t)c * WAVE;
    const uint32_t i = j * span + shift;
    xc[c] = j < used ? seg_x(m, fences[m.r0 + i]) : 0.0f;
    tc[c] = (float)i * inv_n;
  }
  for (uint32_t it = 0; it < steps; ++it) {
    float p[MLP_PARAMS];
    broadcast_params(pv, p);
    float g[MLP_PARAMS];
#pragma unroll
    for (int q = 0; q < MLP_PARAMS; ++q) g[q] = 0.0f;
#pragma unroll
    for (int c = 0; c < TRAIN_CACHE; ++c)
      if ((uint32_t)lane + (uint32_t)c * WAVE < used) mlp_accumulate(p, xc[c], tc[c], scale, g);
    const float gr = wave_reduce_scatter(g);
    b1t *= beta1;
    b2t *= beta2;
    am = beta1 * am + (1.0f - beta1) * gr;
    av = beta2 * av + (1.0f - beta2) * gr * gr;
    const float step_lr = lr * (1.0f - 0.9f * (float)it / (float)steps);
    if (owner) pv = pv - step_lr * (am / (1.0f - b1t)) / (sqrtf(av / (1.0f - b2t)) + eps);
  }
  if (owner) m.p[own] = pv;
}

// Loss and gradient of the mean of 0.5 * (mlp(x_i) - t_i)^2 over n pairs (one wavefront).
__global__ __launch_bounds__(WAVE) void k_mlp_grad(const float* __restrict__ x, const float* __restrict__ t,
                                                   uint32_t n, const float* __restrict__ params,
                                                   float* __restrict__ grad, float* __restrict__ loss) {
  const int own = param_of_lane(lane_id());
  float p[MLP_PARAMS];
  broadcast_params(own < MLP_PARAMS ? params[own] : 0.0f, p);
  float g[MLP_PARAMS];
  for (int q = 0; q < MLP_PARAMS; ++q) g[q] = 0.0f;
  float l = 0.0f;
  const float inv_n = n ? 1.0f / (float)n : 0.0f;
  for (uint32_t i = lane_id(); i < n; i += WAVE) l += mlp_accumulate(p, x[i], t[i], inv_n, g);
  const float gr = wave_reduce_scatter(g);
  if (own < MLP_PARAMS) grad[own] = gr;
  for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off, WAVE);
  if (lane_id() == 0) *loss = l;
}

// Window hit rate of the model on its own training pairs: rank r is a hit when the
// prediction for fences[r] lies within ROUTE_WINDOW of r (the lookup's window then holds r).
__global__ void k_route_eval(const RouteHeader* __restrict__ rh, const uint64_t* __restrict__ fences,
                             const uint32_t* __restrict__ lids, const SegModel* __restrict__ segs,
                             uint32_t n_max, unsigned long long* __restrict__ hm) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const RouteView v = route_view(rh, fences, lids, segs);
  bool valid = r < v.L && r < n_max;
  bool hit = false;
  if (valid) {
    const uint64_t key = fences[r];
    const SegModel& m = segs[root_segment(v, key)];
    float rank = (float)m.r0 + mlp_forward(m.p, seg_x(m, key)) * m.n;
    rank = fminf(fmaxf(rank, 0.0f), (float)(v.L - 1));
    const int64_t p = (int64_t)(rank + 0.5f);
    hit = p >= (int64_t)r - ROUTE_WINDOW && p <= (int64_t)r + ROUTE_WINDOW;
  }
  const uint64_t h = ballot(valid && hit), m = ballot(valid && !hit);
  if (lane_id() == 0) {
    if (h) atomicAdd(hm, (unsigned long long)__popcll(h));
    if (m) atomicAdd(hm + 1, (unsigned long long)__popcll(m));
  }
}

hipError_t route_evaluate(const RouteState& rs, uint64_t* d_hm, uint32_t n_max, hipStream_t s) {
  if (n_max == 0) return hipSuccess;
  k_route_eval<<<(n_max + 255) / 256, 256, 0, s>>>(rs.hdr, rs.fences, rs.lids, rs.segs, n_max,
                                                   (unsigned long long*)d_hm);
  return hipGetLastError();
}

// ------------------------------------------------------------------ host side

hipError_t route_build(const TreeView& t, RouteState& rs, uint32_t n_nodes, uint32_t steps,
                       float lr, bool train, bool warm, hipStream_t s) {
  if (n_nodes == 0 || n_nodes > t.capacity) return hipErrorInvalidValue;
  rs.builds += 1;
  hipError_t e;
  if (warm) {
    // Rewrite the existing directory: leaves at most n_nodes, so every grid is sized by it.
    const uint32_t g = (n_nodes + 255) / 256;
    const uint32_t gt = (n_nodes + DS_TILE - 1) / DS_TILE;
    uint32_t* counts = rs.order;
    uint32_t* offs = rs.order_tmp;
    uint32_t* tiles = sort_spare_words(rs.hist);  // one word per scan tile: n_nodes / DS_TILE
    k_dir_count<<<g, 256, 0, s>>>(t, rs.hdr, rs.dir, counts);
    k_dir_scan_reduce<<<gt, DS_BLOCK, 0, s>>>(rs.hdr, counts, tiles);
    k_dir_scan_tiles<<<1, DS_BLOCK, 0, s>>>(rs.hdr, tiles);
    k_dir_scan_apply<<<gt, DS_BLOCK, 0, s>>>(rs.hdr, counts, tiles, offs);
    k_dir_rewrite<<<g, 256, 0, s>>>(t, rs.hdr, rs.dir, counts, offs, rs.dir_alt, rs.rank_of);
    k_dir_install<<<(n_nodes + 1 + 255) / 256, 256, 0, s>>>(rs.hdr, rs.dir_alt, rs.dir, rs.fences, rs.lids);
    k_dir_set_len<<<1, 64, 0, s>>>(rs.hdr);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else {
    k_route_lows<<<(t.capacity + 255) / 256, 256, 0, s>>>(t, n_nodes, rs.lows, rs.rank_of);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(rs.lows, rs.lows_tmp, rs.sorted_lows, rs.order_tmp, rs.order, n_nodes, 0, 64,
                         rs.hist, s);
    if (e != hipSuccess) return e;
    // Cold builds are rare: check the sort's stall word before its order is used.
    uint32_t stalled = 0;
    if ((e = hipMemcpyAsync(&stalled, sort_stall_word(rs.hist), 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return e;
    if (stalled) return hipErrorLaunchTimeOut;
    k_route_directory<<<(n_nodes + 255) / 256, 256, 0, s>>>(rs.sorted_lows, rs.order, n_nodes, rs.hdr,
                                                            rs.fences, rs.lids, rs.dir, rs.rank_of);
    k_route_parents<<<(n_nodes + 255) / 256, 256, 0, s>>>(t, n_nodes, rs.dir);
  }
  if (!warm) k_route_root<<<1, 64, 0, s>>>(rs.fences, rs.hdr);
  k_route_segments<<<(ROUTE_SEGMENTS + 255) / 256, 256, 0, s>>>(rs.hdr, rs.fences, rs.lids, rs.segs,
                                                                warm ? 0u : 1u);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (train && steps > 0) {
    k_route_train<<<(ROUTE_SEGMENTS + TRAIN_WAVES - 1) / TRAIN_WAVES, TRAIN_WAVES * WAVE, 0, s>>>(
        rs.hdr, rs.fences, rs.segs, steps, lr, rs.builds);
    e = hipGetLastError();
  }
  return e;
}

hipError_t mlp_grad(const float* d_x, const float* d_t, uint32_t n, const float* d_params,
                    float* d_grad, float* d_loss, hipStream_t s) {
  k_mlp_grad<<<1, WAVE, 0, s>>>(d_x, d_t, n, d_params, d_grad, d_loss);
  return hipGetLastError();
}

}  // namespace stm
