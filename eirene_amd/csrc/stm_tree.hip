// stm_tree.hip — B+tree kernels for gfx950: bottom-up build, read-only batched lookup,
// transactional upsert (STM), dump and structural check.
//
// Reference counterpart: tx_insert_kernel (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:15-258)
// with its device library (bpt-gpustm-kernel-tree.cuh, bpt-gpustm-kernel-tx.cuh).  The
// reference runs every unique key of a batch as a transaction on one 32-lane warp.  Here
// the merged ops of a batch are split by type (the merged ops touch distinct keys, so
// they commute):
//   * merged GETs run first in k_lookup on a tree no one writes, walking the sorted keys
//     in chunks with the last inner node and leaf cached in registers;
//   * merged INSERTs then run in k_upsert as transactions on 64-lane wavefronts:
//     begin (root + version), optimistic reads validated by fences and lock versions,
//     try-lock commit of the write set with 64-bit CAS, abort and retry with backoff,
//     preemptive split of full inner nodes on the way down
//     (reference: src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:162-181).
#include <atomic>
#include "stm_common.hpp"
#include "stm_internal.hpp"
#include "stm_learned.hpp"
#include "stm_tree.hpp"

namespace stm {

constexpr uint32_t LOOKUP_GRID_CAP = 8192;

// ================================================================== bulk build
template <int F>
__global__ void k_build_leaves(TreeView t, const uint64_t* __restrict__ ukey,
                               const uint64_t* __restrict__ uval, uint32_t U, uint32_t fill,
                               uint32_t L, uint64_t* __restrict__ lows_out) {
  const uint64_t gi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gi >= (uint64_t)L * F) return;
  const uint32_t leaf = (uint32_t)(gi / F);
  const uint32_t s = (uint32_t)(gi % F);
  const uint64_t first = (uint64_t)leaf * fill;
  const uint64_t src = first + s;
  const bool valid = s < fill && src < U;
  t.keys[gi] = valid ? ukey[src] : KEY_EMPTY;
  t.vals[gi] = valid ? uval[src] : 0;
  if (s == 0) {
    const uint32_t cnt = U > first ? (uint32_t)min<uint64_t>(fill, U - first) : 0u;
    NodeMeta m{};
    m.lock = 0;
    m.info = make_info(cnt, 0);
    m.low = leaf == 0 ? 0 : ukey[first];
    m.high = (leaf + 1 < L) ? ukey[first + fill] : KEY_EMPTY;
    m.next = (leaf + 1 < L) ? (uint64_t)(leaf + 1) : (uint64_t)NODE_NONE;
    t.meta[leaf] = m;
    lows_out[leaf] = m.low;
  }
}

template <int F>
__global__ void k_build_inner(TreeView t, const uint64_t* __restrict__ lows_prev, uint32_t n_prev,
                              uint32_t first_prev, uint32_t fill, uint32_t n_this,
                              uint32_t first_this, uint32_t level, uint64_t* __restrict__ lows_out) {
  const uint64_t gi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gi >= (uint64_t)n_this * F) return;
  const uint32_t j = (uint32_t)(gi / F);
  const uint32_t s = (uint32_t)(gi % F);
  const uint64_t c = (uint64_t)j * fill + s;
  const bool valid = s < fill && c < n_prev;
  const uint64_t node = (uint64_t)first_this + j;
  t.keys[node * F + s] = valid ? lows_prev[c] : KEY_EMPTY;
  t.vals[node * F + s] = valid ? (uint64_t)(first_prev + c) : 0;
  if (s == 0) {
    const uint64_t c0 = (uint64_t)j * fill;
    NodeMeta m{};
    m.lock = 0;
    m.info = make_info((uint32_t)min<uint64_t>(fill, n_prev - c0), level);
    m.low = lows_prev[c0];
    m.high = (j + 1 < n_this) ? lows_prev[c0 + fill] : KEY_EMPTY;
    m.next = (j + 1 < n_this) ? (uint64_t)(node + 1) : (uint64_t)NODE_NONE;
    t.meta[node] = m;
    lows_out[j] = m.low;
  }
}

// ================================================================== read-only lookup
// G lanes serve one op; lane g of the group holds slots [g*K, g*K+K) of a node.
template <int G>
STM_DEV uint32_t group_sum(uint32_t v) {
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// x[e] for a runtime e, written as a compile-time select tree so the array stays in
// registers (a loop here lets the compiler turn it into an indexed scratch access).
template <int K, int I = 0>
STM_DEV uint64_t select_k(const uint64_t (&x)[K], int e) {
  if constexpr (I == K - 1) {
    return x[I];
  } else {
    return e == I ? x[I] : select_k<K, I + 1>(x, e);
  }
}

// Slot s of the node held by the group (s group-uniform, 0 <= s < F).
template <int F, int G>
STM_DEV uint64_t group_at(const uint64_t (&x)[F / G], int s, int gbase) {
  constexpr int K = F / G;
  return shfl64(select_k<K>(x, s % K), gbase + s / K);
}

template <int F, int G>
STM_DEV int group_count_le(const uint64_t (&x)[F / G], uint64_t key) {
  constexpr int K = F / G;
  uint32_t c = 0;
#pragma unroll
  for (int e = 0; e < K; ++e) c += (x[e] <= key) ? 1u : 0u;
  return (int)group_sum<G>(c);
}

// Slot of key in the leaf held by the group (lane g holds slots g*K .. g*K+K-1, as load_slots
// loads them), -1 if absent.  A leaf's entries are unordered and its free slots hold KEY_EMPTY,
// which is never a key, so one equality test per slot finds it.
template <int F, int G>
STM_DEV int group_find(const uint64_t (&x)[F / G], uint64_t key, int g) {
  constexpr int K = F / G;
  uint32_t hit = 0;  // e + 1 of the slot that holds key in this lane
#pragma unroll
  for (int e = 0; e < K; ++e) hit = x[e] == key ? (uint32_t)(e + 1) : hit;
  return (int)group_sum<G>(hit ? (uint32_t)(g * K) + hit : 0u) - 1;
}

template <int F, int G>
STM_DEV void load_slots(const uint64_t* __restrict__ base, uint32_t node, int g,
                        uint64_t (&x)[F / G]) {
  constexpr int K = F / G;
  const uint64_t* p = base + (size_t)node * F + (size_t)g * K;
#pragma unroll
  for (int e = 0; e < K; ++e) x[e] = p[e];
}

template <int G>
STM_DEV float group_sumf(float v) {
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// Per lane-group state of the learned path: a window of G consecutive directory entries,
// entry base + g in lane g (fence, leaf id, split flag).  Keys arrive sorted, so most keys
// route inside the window with no memory access; the window is refilled from the next
// entries when a key passes its end, and from the model's prediction otherwise.
struct RouteCursor {
  bool valid = false;
  uint32_t base = 0;
  uint64_t f = KEY_EMPTY;  // fence of entry base + g (KEY_EMPTY past the last entry)
  uint32_t lid = NODE_NONE;
  uint32_t pflags = DIR_SPLIT | DIR_PAR_NONE;
};

template <int G>
STM_DEV void window_load(const RouteView& rv, uint32_t base, int g, RouteCursor& c) {
  const uint64_t idx = (uint64_t)base + (uint32_t)g;
  c.valid = true;
  c.base = base;
  c.f = KEY_EMPTY;
  c.lid = NODE_NONE;
  c.pflags = DIR_SPLIT | DIR_PAR_NONE;
  if (idx < rv.L) {
    const DirEntry e = rv.dir[idx];
    c.f = e.fence;
    c.lid = e.lid;
    c.pflags = e.pflags;
  }
}

// Entry of `key` inside the window: true when some entry base + j holds it with its
// successor's fence also in the window (j <= G - 2).
template <int G>
STM_DEV bool window_find(const RouteCursor& c, uint64_t key, int gbase, uint32_t& lid, bool& dirty,
                         uint64_t& hi, uint32_t& cnt_out, uint32_t& par) {
  const uint32_t cnt = group_sum<G>(c.valid && c.f <= key ? 1u : 0u);
  cnt_out = cnt;
  const uint32_t j = cnt > 0 ? cnt - 1 : 0;
  lid = (uint32_t)__shfl((int)c.lid, gbase + (int)j, WAVE);
  const uint32_t pf = (uint32_t)__shfl((int)c.pflags, gbase + (int)j, WAVE);
  dirty = (pf & DIR_SPLIT) != 0;
  par = (pf & DIR_PAR_NONE) == DIR_PAR_NONE ? NODE_NONE : (pf & DIR_PAR_NONE);
  hi = shfl64(c.f, gbase + (int)min(cnt, (uint32_t)G - 1));
  return c.valid && cnt >= 1 && cnt <= (uint32_t)G - 1;
}

// Learned routing of `key` to a directory entry on a tree nobody writes.  Returns true with
// the entry's leaf id, split flag and upper fence; false sends the caller to a root-to-leaf
// descent.  model: 0 = routed without the model, 1 = the model's +/-2 window held the key,
// 2 = the model missed.
template <int F, int G, bool INDEX>
STM_DEV bool learned_route(const RouteView& rv, uint64_t key, int g, int gbase, RouteCursor& c,
                           uint32_t& lid, bool& dirty, uint64_t& hi, int& model, uint32_t& par) {
  model = 0;
  if (rv.L == 0) return false;
  uint32_t cnt = 0;
  if (window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par)) return true;
  if (c.valid && cnt == (uint32_t)G && c.base + G - 1 < rv.L) {
    // Past the window: the next entries (one coalesced load) usually hold the key.
    window_load<G>(rv, c.base + G - 1, g, c);
    if (window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par)) return true;
  }
  if constexpr (INDEX) {
    // Static index: a G-ary search of the directory for the last fence <= key.  The rank lies
    // in [lo, top); each step every lane of the group probes one fence, and the range shrinks
    // to one probe interval.  A range of at most G - 1 entries fits a window with the
    // successor of its last entry.
    uint32_t lo = 0, top = rv.L;
    while (top - lo > (uint32_t)G - 1) {
      const uint32_t step = (top - lo + G - 1) / G;
      const uint32_t q = lo + (uint32_t)g * step;
      const uint32_t below = group_sum<G>(q < top && rv.fences[q] <= key ? 1u : 0u);
      if (below == 0) break;  // the key precedes the first fence: descend instead
      lo += (below - 1) * step;
      top = min(top, lo + step);
    }
    window_load<G>(rv, lo, g, c);
    const bool ok = window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par);
    model = ok ? 1 : 2;
    return ok;
  }
  // The model: root line -> segment -> MLP -> rank p; entries [p-2, p-2+G) are loaded and the
  // key must lie in [p-2, p+2] for a model hit.
  const SegModel* m = rv.segs + root_segment(rv, key);
  const float x = (float)(((double)key - (double)m->klo) * (double)m->inv_w);
  float acc = 0.0f;
  for (int j = g; j < MLP_HIDDEN; j += G)
    acc += m->p[2 * MLP_HIDDEN + j] * fmaxf(m->p[j] * x + m->p[MLP_HIDDEN + j], 0.0f);
  const float y = group_sumf<G>(acc) + m->p[3 * MLP_HIDDEN];
  float rank = (float)m->r0 + y * m->n;
  rank = fminf(fmaxf(rank, 0.0f), (float)(rv.L - 1));  // also maps NaN to 0
  const int64_t p = (int64_t)(rank + 0.5f);
  const uint32_t base = (uint32_t)max<int64_t>(p - ROUTE_WINDOW, 0);
  window_load<G>(rv, base, g, c);
  bool ok = window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par);
  const int64_t r = (int64_t)base + (int64_t)cnt - 1;
  model = (ok && r >= p - ROUTE_WINDOW && r <= p + ROUTE_WINDOW) ? 1 : 2;
  if (!ok) {
    // A miss is usually near: one more window on the side the key lies (before the window:
    // no entry <= key; past it: every entry <= key) before the caller's root-to-leaf descent.
    if (cnt == 0 && base > 0) {
      window_load<G>(rv, base > (uint32_t)G - 1 ? base - (uint32_t)(G - 1) : 0u, g, c);
      ok = window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par);
    } else if (cnt == (uint32_t)G && base + G - 1 < rv.L) {
      window_load<G>(rv, base + G - 1, g, c);
      ok = window_find<G>(c, key, gbase, lid, dirty, hi, cnt, par);
    }
  }
  return ok;
}

// Stripe of the statistics counters a wavefront adds to.
STM_DEV uint32_t wave_stripe() { return blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; }

// Latency-bound: resident wavefronts hide the node loads.  For F <= 64 and G >= 16 the
// register budget of seven wavefronts per SIMD (72 VGPRs) is requested; the learned variant
// otherwise lands one register above it and loses a seventh of its wavefronts.  Narrower lane
// groups hold more slots per lane and would spill.
template <int F, int G>
constexpr int lookup_min_waves() { return (F <= 64 && G >= 16) ? 7 : 1; }

template <int F, int G, bool LEARNED, bool INDEX>
__global__ __launch_bounds__(256, (lookup_min_waves<F, G>())) void k_lookup(TreeView t, const uint64_t* __restrict__ ukey,
                                                const uint8_t* __restrict__ uop,
                                                const uint32_t* __restrict__ gstart,
                                                const uint32_t* __restrict__ npend,
                                                const uint32_t* __restrict__ sidx,
                                                const BatchCounters* __restrict__ cnt,
                                                uint32_t chunk, uint64_t* __restrict__ out_vals,
                                                uint8_t* __restrict__ out_found,
                                                const RouteHeader* __restrict__ rh,
                                                const uint64_t* __restrict__ fences,
                                                const uint32_t* __restrict__ lids,
                                                const SegModel* __restrict__ segs,
                                                const BatchDesc* __restrict__ desc,
                                                InsHint* __restrict__ ihint,
                                                uint64_t* __restrict__ gval,
                                                uint64_t* __restrict__ fanout,
                                                uint64_t fanout_cap) {
  constexpr int K = F / G;
  const int lane = lane_id();
  const int g = lane & (G - 1);
  const int gbase = lane - g;
  const uint64_t grp0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
  const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / G;
  const uint32_t U = cnt->n_unique;
  // Every lane group walks one contiguous span of at least `chunk` sorted groups, so its
  // cached leaf, parent and directory cursor carry over for the whole span.
  const uint64_t span = max<uint64_t>(chunk, (U + ngroups - 1) / ngroups);
  // Groups past the end skip the walk but stay for the wave's statistics reduction.
  const bool active = grp0 * span < U;
  if (desc) {
    out_vals = desc->out_vals;
    out_found = desc->out_found;
  }
  const uint32_t cap = t.capacity;
  const uint32_t root = (uint32_t)min<uint64_t>(t.hdr->root, cap - 1);
  const uint32_t height = (uint32_t)min<uint64_t>(t.hdr->height, 64);

  uint32_t leaf = NODE_NONE;
  uint64_t leaf_hi = 0;
  uint64_t lk[K];
  uint32_t par = NODE_NONE;
  uint64_t par_hi = 0;
  uint64_t pk[K], pv[K];
  RouteView rv{};
  RouteCursor cur;
  uint64_t cur_hi = 0;  // upper fence of the current directory entry
  uint32_t hits = 0, misses = 0;
  uint32_t reads = 0;  // tree nodes this lane group loaded
  if constexpr (LEARNED) rv = route_view(rh, fences, lids, segs, t.dir);

  if (active) {
  const uint32_t i0 = (uint32_t)(grp0 * span);
  const uint32_t i1 = (uint32_t)min<uint64_t>((uint64_t)i0 + span, U);
  for (uint32_t base = i0; base < i1; base += G) {
    // Lane g owns op base + g: its metadata comes in one coalesced round trip, and the
    // sorted position of its first waiting op in a second one that overlaps the walk.
    const uint32_t j = base + g;
    const bool valid = j < i1;
    // Loads from an index clamped into the span (i1 > base), selected afterwards: unconditional
    // loads issue together instead of one guarded round trip after another.
    const uint32_t jc = valid ? j : i1 - 1;
    const uint8_t op_raw = uop[jc];
    const uint64_t key_raw = ukey[jc];
    const uint32_t gs_raw = gstart[jc];
    const uint32_t np_raw = npend[jc];
    __builtin_amdgcn_sched_barrier(0);  // keep all four loads ahead of their first use
    const uint8_t op_l = valid ? op_raw : UOP_SKIP;
    const uint64_t key_l = valid ? key_raw : 0;
    const uint32_t gs_l = valid ? gs_raw : 0;
    const uint32_t np_l = valid ? np_raw : 0;
    const bool mine = op_l == OP_GET;
    // Merged INSERTs are routed too: their leaf is the upsert kernel's starting hint.
    const bool walk = mine || (ihint && op_l == OP_INSERT);
    const uint32_t sx_raw = sidx[gs_l];  // gs_l is a sorted position of the batch
    const uint32_t pos_l = (mine && np_l) ? pos_of(sx_raw) : 0;
    uint64_t vidx_l = 0;  // value slot of op base + g
    uint32_t leaf_l = NODE_NONE, par_l = NODE_NONE;
    bool found_l = false;
    const uint32_t n_sub = min(i1 - base, (uint32_t)G);
    for (uint32_t jj = 0; jj < n_sub; ++jj) {
      if (__shfl((int)walk, gbase + (int)jj, WAVE) == 0) continue;
      const uint64_t key = shfl64(key_l, gbase + (int)jj);
      if (leaf == NODE_NONE || key >= leaf_hi) {
        bool routed = false;
        if constexpr (LEARNED) {
          uint32_t lid = leaf;
          bool dirty = true;
          int model = 0;
          uint32_t rpar = par;
          // Still inside the current directory entry: the leaf split, walk right from it.
          const bool same_entry = leaf != NODE_NONE && key < cur_hi;
          uint64_t hi = cur_hi;
          routed = same_entry ||
                   learned_route<F, G, INDEX>(rv, key, g, gbase, cur, lid, dirty, hi, model, rpar);
          if (routed) {
            par = rpar;  // the level-1 hint recorded for merged INSERTs
            cur_hi = hi;
            leaf = min(lid, cap - 1);
            uint64_t mh = KEY_EMPTY;
            if (dirty) mh = t.meta[leaf].high;
            load_slots<F, G>(t.keys, leaf, g, lk);
            reads += 1;
            if (dirty && key >= mh) {  // the key lives in a right sibling made by a split
              routed = false;
              uint32_t steps = 0;
              for (int step = 0; step < 64; ++step) {
                leaf = (uint32_t)min<uint64_t>(t.meta[leaf].next, cap - 1);
                mh = t.meta[leaf].high;
                steps = (uint32_t)step + 1;
                if (key < mh) {
                  routed = true;
                  break;
                }
              }
              // Rare (only after splits): counted where it happens, so no register carries it.
              if (g == 0) stat_add(t.stats, ST_LEAF_STEPS, blockIdx.x, steps);
              if (routed) {
                load_slots<F, G>(t.keys, leaf, g, lk);
                reads += 1;
              }
              else cur_hi = 0;
            }
            leaf_hi = min(cur_hi, mh);
          } else {
            cur_hi = 0;
          }
          hits += model == 1 ? 1 : 0;
          misses += model == 2 ? 1 : 0;
        }
        if (!routed) {
          if (height <= 1) {
            leaf = root;
            leaf_hi = KEY_EMPTY;
          } else {
            // The learned variant keeps no cached parent: its misses are rare and the
            // registers buy occupancy.
            if (LEARNED || par == NODE_NONE || key >= par_hi) {
              uint32_t node = root;
              uint64_t hi = KEY_EMPTY;
              for (uint32_t lvl = height - 1; lvl > 1; --lvl) {
                uint64_t nk[K], nv[K];
                load_slots<F, G>(t.keys, node, g, nk);
                load_slots<F, G>(t.vals, node, g, nv);
                reads += 1;
                const int s = max(group_count_le<F, G>(nk, key) - 1, 0);
                const uint64_t child = group_at<F, G>(nv, s, gbase);
                const uint64_t nx = (s + 1 < F) ? group_at<F, G>(nk, s + 1, gbase) : KEY_EMPTY;
                hi = min(hi, nx);
                node = (uint32_t)min<uint64_t>(child, cap - 1);
              }
              load_slots<F, G>(t.keys, node, g, pk);
              load_slots<F, G>(t.vals, node, g, pv);
              reads += 1;
              par = node;
              par_hi = hi;
            }
            const int s = max(group_count_le<F, G>(pk, key) - 1, 0);
            leaf = (uint32_t)min<uint64_t>(group_at<F, G>(pv, s, gbase), cap - 1);
            const uint64_t nx = (s + 1 < F) ? group_at<F, G>(pk, s + 1, gbase) : KEY_EMPTY;
            leaf_hi = min(par_hi, nx);
          }
          load_slots<F, G>(t.keys, leaf, g, lk);
          reads += 1;
        }
      }
      const int s = group_find<F, G>(lk, key, g);
      if (g == (int)jj) {
        found_l = s >= 0;
        // The key's slot when found (a new key is appended at the leaf's count).
        vidx_l = (uint64_t)leaf * F + (uint32_t)(found_l ? s : 0);
        leaf_l = leaf;
        par_l = par;
      }
    }
    // The sub-chunk's tail loads go out together, before any of its stores (the vector memory
    // counter counts stores too, so a load issued after a store also waits for the store):
    // the version and entry count of an INSERT group's leaf, the value a GET group returns.
    const bool hint = walk && !mine;
    uint64_t ver = 1ull, info = 0;
    if (hint && leaf_l < t.capacity) {
      ver = t.meta[leaf_l].lock;
      info = t.meta[leaf_l].info;
    }
    const uint64_t val = (mine && found_l) ? t.vals[vidx_l] : 0;
    if (hint) {
      // The record for the upsert kernels: the op's answer range and leaf, the leaf's version
      // and entry count now (no writer runs during this kernel), the key's slot in it and the
      // level-1 node last used for routing.
      // field by field, so that the fields need not be live together (registers)
      InsHint* h = ihint + j;
      h->ver = ver;
      h->leaf = leaf_l;
      h->pos = (uint32_t)(vidx_l - (uint64_t)leaf_l * F) | (found_l ? UPOS_FOUND : 0u) |
               (min(meta_count(info), 255u) << UPOS_COUNT_SHIFT);
      h->par = par_l;
      h->gs = gs_l;
      h->np = np_l;
    }
    // The answers.
    if (mine) {
      if (np_l > FANOUT_DIRECT) {
        // A hot key: its waiting ops are answered by the whole grid (fanout_answers), in
        // chunks queued here, instead of one store after another from this lane.
        gval[j] = val;
        const uint32_t nch = (np_l + FANOUT_CHUNK - 1) / FANOUT_CHUNK;
        const uint32_t c0 = atomicAdd(&const_cast<BatchCounters*>(cnt)->n_fanout, nch);
        for (uint32_t c = 0; c < nch; ++c)
          if ((uint64_t)c0 + c < fanout_cap)
            fanout[c0 + c] = (uint64_t)j | ((uint64_t)(c * FANOUT_CHUNK) << 32) |
                             (found_l ? (1ull << 63) : 0ull);
      } else {
        if (np_l) {
          out_vals[pos_l] = val;
          out_found[pos_l] = found_l ? 1 : 0;
        }
        for (uint32_t p = 1; p < np_l; ++p) {
          const uint32_t pos = pos_of(sidx[gs_l + p]);
          out_vals[pos] = val;
          out_found[pos] = found_l ? 1 : 0;
        }
      }
    }
  }
  }  // span
  // One atomic per wavefront and counter: the group leaders' counts summed by readlane.
  uint32_t h = 0, m = 0, nr = 0;
#pragma unroll
  for (int q = 0; q < WAVE; q += G) {
    if constexpr (LEARNED) {
      h += (uint32_t)__builtin_amdgcn_readlane((int)hits, q);
      m += (uint32_t)__builtin_amdgcn_readlane((int)misses, q);
    }
    nr += (uint32_t)__builtin_amdgcn_readlane((int)reads, q);
  }
  if (lane == 0) {
    if (h) stat_add(t.stats, ST_ROUTE_HITS, wave_stripe(), h);
    if (m) stat_add(t.stats, ST_ROUTE_MISSES, wave_stripe(), m);
    if (nr) stat_add(t.stats, ST_NODE_READS, wave_stripe(), nr);
  }
}

// ================================================================== transactional upsert
// A node held by one wavefront: slot s = j*64 + lane.
template <int F>
struct WNode {
  static constexpr int SPL = (F + WAVE - 1) / WAVE;
  uint64_t k[SPL];
  uint64_t v[SPL];
  uint64_t ver;
  uint64_t low, high;
  uint32_t id, count, level, next;
};

template <int F>
STM_DEV uint64_t wave_at(const uint64_t (&x)[WNode<F>::SPL], uint32_t s) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j)
    if ((s >> 6) == (uint32_t)j) r = readlane64(x[j], (int)(s & 63));
  return r;
}

// Per-lane gather of slot src (any lane, any j); dflt when src is outside [0, F).
template <int F>
STM_DEV uint64_t wgather(const uint64_t (&x)[WNode<F>::SPL], int src, uint64_t dflt) {
  uint64_t r = dflt;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const uint64_t tv = shfl64(x[j], src & 63);
    if (src >= 0 && src < F && (src >> 6) == j) r = tv;
  }
  return r;
}

template <int F>
STM_DEV uint32_t wave_count_le(const uint64_t (&x)[WNode<F>::SPL], uint64_t key) {
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) c += (uint32_t)__popcll(ballot(x[j] <= key));
  return c;
}

// Slot of key in a leaf held by the wavefront (entries unordered, free slots KEY_EMPTY, which
// is never a key); NODE_NONE if absent.  Wave-uniform.
template <int F>
STM_DEV uint32_t wave_find(const uint64_t (&x)[WNode<F>::SPL], uint64_t key) {
  uint32_t s = NODE_NONE;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const uint64_t m = ballot(x[j] == key);
    if (m) s = (uint32_t)(j * WAVE) + (uint32_t)__builtin_ctzll(m);
  }
  return s;
}

// Rank of each held key among the leaf's first `count` slots (the number of those keys below
// it): its place in key order.  The keys are distinct, so the ranks of the count entries are a
// permutation of [0, count).  One wave-uniform pass over the entries.
template <int F>
STM_DEV void wave_ranks(const uint64_t (&x)[WNode<F>::SPL], uint32_t count,
                        uint32_t (&r)[WNode<F>::SPL]) {
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) r[j] = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const uint64_t ki = wave_at<F>(x, i);
#pragma unroll
    for (int j = 0; j < WNode<F>::SPL; ++j) r[j] += ki < x[j] ? 1u : 0u;
  }
}

// Number of the leaf's first `count` entries below key (wave-uniform).
template <int F>
STM_DEV uint32_t wave_count_below(const uint64_t (&x)[WNode<F>::SPL], uint32_t count, uint64_t key) {
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j)
    c += (uint32_t)__popcll(ballot((uint32_t)(j * WAVE + lane_id()) < count && x[j] < key));
  return c;
}

// Slot of parent p that routes key, checked to point at child; NODE_NONE if it does not.
template <int F>
STM_DEV uint32_t parent_slot(const WNode<F>& p, uint64_t key, uint32_t child) {
  const uint32_t cle = wave_count_le<F>(p.k, key);
  if (cle == 0 || cle > p.count) return NODE_NONE;
  if (wave_at<F>(p.v, cle - 1) != (uint64_t)child) return NODE_NONE;
  return cle - 1;
}

// Optimistic read of a node: version word, then (ordered after it) the content.  The
// content is only trusted after the version is re-validated: by the commit CAS for nodes
// the transaction writes, by fences and levels for the nodes it only passes through.
template <int F>
STM_DEV bool tx_read(const TreeView& t, uint32_t id, WNode<F>& n) {
  if (id >= t.capacity) return false;
  NodeMeta* m = t.meta + id;
  const uint64_t w = uniform64(ld_relaxed(&m->lock));
  if (w & 1) return false;
  wave_order();  // content loads issue only after the version word has been read
  const uint64_t info = uniform64(ld_relaxed(&m->info));
  n.low = uniform64(ld_relaxed(&m->low));
  n.high = uniform64(ld_relaxed(&m->high));
  n.next = (uint32_t)uniform64(ld_relaxed(&m->next));
  const uint64_t* kp = t.keys + (size_t)id * F;
  const uint64_t* vp = t.vals + (size_t)id * F;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const int s = j * WAVE + lane_id();
    if (s < F) {
      n.k[j] = ld_relaxed(kp + s);
      n.v[j] = ld_relaxed(vp + s);
    } else {
      n.k[j] = KEY_EMPTY;
      n.v[j] = 0;
    }
  }
  n.ver = w;
  n.id = id;
  n.count = meta_count(info);
  n.level = meta_level(info);
  return n.count <= (uint32_t)F;
}

// The metadata words of node id (count | level, fences, right sibling), one lane each.
STM_DEV void tx_write_meta(const TreeView& t, uint32_t id, uint32_t count, uint32_t level,
                           uint64_t low, uint64_t high, uint32_t next) {
  NodeMeta* m = t.meta + id;
  const int l = lane_id();
  if (l == 1) st_relaxed(&m->info, make_info(count, level));
  if (l == 2) st_relaxed(&m->low, low);
  if (l == 3) st_relaxed(&m->high, high);
  if (l == 4) st_relaxed(&m->next, (uint64_t)next);
}

template <int F>
STM_DEV void tx_write_node(const TreeView& t, uint32_t id, const uint64_t (&k)[WNode<F>::SPL],
                           const uint64_t (&v)[WNode<F>::SPL], uint32_t count, uint32_t level,
                           uint64_t low, uint64_t high, uint32_t next) {
  uint64_t* kp = t.keys + (size_t)id * F;
  uint64_t* vp = t.vals + (size_t)id * F;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const int s = j * WAVE + lane_id();
    if (s < F) {
      st_relaxed(kp + s, k[j]);
      st_relaxed(vp + s, v[j]);
    }
  }
  tx_write_meta(t, id, count, level, low, high, next);
}

// Try-lock of up to three version words (two nodes and optionally the tree header).
// Each word must still hold the version the transaction read.  On any failure the words
// already taken are restored to their old value (nothing was written) and false returns.
STM_DEV bool tx_lock(uint64_t* w0, uint64_t v0, uint64_t* w1, uint64_t v1, uint64_t* w2,
                     uint64_t v2) {
  const int l = lane_id();
  uint64_t* w = l == 0 ? w0 : (l == 1 ? w1 : (l == 2 ? w2 : nullptr));
  const uint64_t v = l == 0 ? v0 : (l == 1 ? v1 : v2);
  bool ok = true;
  if (w) ok = cas_u64(w, v, v | 1ull);
  const uint64_t fail = ballot(!ok);
  if (fail) {
    if (w && ok) st_relaxed(w, v);
    return false;
  }
  wave_order();  // the lock bits are held before any content store issues
  return true;
}

STM_DEV void tx_unlock(uint64_t* w0, uint64_t v0, uint64_t* w1, uint64_t v1, uint64_t* w2,
                       uint64_t v2) {
  wave_order();  // every content store is acknowledged before the new versions
  const int l = lane_id();
  uint64_t* w = l == 0 ? w0 : (l == 1 ? w1 : (l == 2 ? w2 : nullptr));
  const uint64_t v = l == 0 ? v0 : (l == 1 ? v1 : v2);
  if (w) st_relaxed(w, v + 2);  // clears the lock bit and bumps the version
}

// Allocates n fresh node ids; NODE_NONE when the pool is exhausted.
STM_DEV uint32_t tx_alloc(const TreeView& t, uint32_t n) {
  uint64_t base = 0;
  if (lane_id() == 0) base = atomicAdd((unsigned long long*)&t.hdr->n_nodes, (unsigned long long)n);
  base = readlane64(base, 0);
  if (base + n > t.capacity) {
    if (lane_id() == 0) atomicOr((unsigned long long*)&t.hdr->error, (unsigned long long)ERR_POOL_FULL);
    return NODE_NONE;
  }
  return (uint32_t)base;
}

// Marks fresh nodes locked (version 0) before they become reachable.
STM_DEV void tx_claim_new(const TreeView& t, uint32_t a, uint32_t b) {
  const int l = lane_id();
  if (l == 0 && a != NODE_NONE) st_relaxed(&t.meta[a].lock, 1ull);
  if (l == 1 && b != NODE_NONE) st_relaxed(&t.meta[b].lock, 1ull);
  wave_order();
}

enum TxResult { TX_DONE = 0, TX_ABORT = 1, TX_RESTART = 2, TX_FAIL = 3, TX_NEED_PARENT = 4 };

struct TxCounters {
  uint32_t commits = 0, aborts = 0, leaf_splits = 0, inner_splits = 0, root_splits = 0;
  uint32_t new_keys = 0;
  uint32_t node_reads = 0;  // nodes read (inner nodes on the way, the leaf)
  bool split_happened = false;  // set by a committed leaf split (the parent changed)
};

// Inserts separator (sep -> rid) right after slot ps of parent p (which has room).
template <int F>
STM_DEV void parent_with_child(const WNode<F>& p, uint32_t ps, uint64_t sep, uint32_t rid,
                               uint64_t (&nk)[WNode<F>::SPL], uint64_t (&nv)[WNode<F>::SPL]) {
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const int s = j * WAVE + lane_id();
    const uint64_t pk = wgather<F>(p.k, s - 1, KEY_EMPTY);
    const uint64_t pv = wgather<F>(p.v, s - 1, 0);
    const bool keep = s <= (int)ps;
    const bool mine = s == (int)ps + 1;
    nk[j] = keep ? p.k[j] : (mine ? sep : pk);
    nv[j] = keep ? p.v[j] : (mine ? (uint64_t)rid : pv);
  }
}

// Splits the full inner node n (preemptively, its own transaction).  pn is its parent when
// has_par; otherwise n is the root, a new root is made and the header is written.
template <int F>
STM_DEV TxResult tx_split_inner(const TreeView& t, const WNode<F>& n, const WNode<F>& pn,
                                bool has_par, uint64_t hver, uint64_t key, TxCounters& c) {
  constexpr int SPL = WNode<F>::SPL;
  const uint32_t h = F / 2;
  const uint64_t sep = wave_at<F>(n.k, h);
  uint32_t ps = 0;
  if (has_par) {
    if (pn.count >= (uint32_t)F) return TX_ABORT;  // parent filled up meanwhile
    ps = parent_slot<F>(pn, key, n.id);
    if (ps == NODE_NONE) return TX_ABORT;
  }
  uint64_t* hw = has_par ? nullptr : &t.hdr->lock;
  if (!tx_lock(&t.meta[n.id].lock, n.ver, has_par ? &t.meta[pn.id].lock : nullptr,
               has_par ? pn.ver : 0, hw, hver))
    return TX_ABORT;
  const uint32_t nnew = has_par ? 1u : 2u;
  const uint32_t base = tx_alloc(t, nnew);
  if (base == NODE_NONE) {
    tx_unlock(&t.meta[n.id].lock, n.ver - 2, has_par ? &t.meta[pn.id].lock : nullptr,
              has_par ? pn.ver - 2 : 0, hw, hver - 2);
    return TX_FAIL;
  }
  const uint32_t rid = base;
  const uint32_t qid = has_par ? NODE_NONE : base + 1;
  tx_claim_new(t, rid, qid);
  uint64_t lk[SPL], lv[SPL], rk[SPL], rv[SPL];
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const int s = j * WAVE + lane_id();
    lk[j] = s < (int)h ? n.k[j] : KEY_EMPTY;
    lv[j] = s < (int)h ? n.v[j] : 0;
    const uint64_t gk = wgather<F>(n.k, s + (int)h, KEY_EMPTY);
    const uint64_t gv = wgather<F>(n.v, s + (int)h, 0);
    const bool r_ok = s < (int)(n.count - h);
    rk[j] = r_ok ? gk : KEY_EMPTY;
    rv[j] = r_ok ? gv : 0;
  }
  tx_write_node<F>(t, rid, rk, rv, n.count - h, n.level, sep, n.high, n.next);
  tx_write_node<F>(t, n.id, lk, lv, h, n.level, n.low, sep, rid);
  if (has_par) {
    uint64_t pk[SPL], pv[SPL];
    parent_with_child<F>(pn, ps, sep, rid, pk, pv);
    tx_write_node<F>(t, pn.id, pk, pv, pn.count + 1, pn.level, pn.low, pn.high,
                     pn.next);
  } else {
    uint64_t qk[SPL], qv[SPL];
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const int s = j * WAVE + lane_id();
      qk[j] = s == 0 ? n.low : (s == 1 ? sep : KEY_EMPTY);
      qv[j] = s == 0 ? (uint64_t)n.id : (s == 1 ? (uint64_t)rid : 0);
    }
    tx_write_node<F>(t, qid, qk, qv, 2, n.level + 1, n.low, n.high, NODE_NONE);
    if (lane_id() == 0) {
      st_relaxed(&t.hdr->root, (uint64_t)qid);
      st_relaxed(&t.hdr->height, (uint64_t)(n.level + 2));
    }
  }
  tx_unlock(&t.meta[n.id].lock, n.ver, has_par ? &t.meta[pn.id].lock : &t.meta[rid].lock,
            has_par ? pn.ver : 0ull, has_par ? &t.meta[rid].lock : &t.hdr->lock, has_par ? 0ull : hver);
  if (!has_par) {
    if (lane_id() == 0) st_relaxed(&t.meta[qid].lock, 2ull);
    c.root_splits++;
  }
  c.inner_splits++;
  c.commits++;
  return TX_RESTART;
}

// Test-and-set lock of a leaf, then its content.  A 64-bit fetch_or takes the lock bit
// without knowing the version; the word it returns is the unlocked version.  Holding the
// lock, the content cannot change, so no re-validation is needed.  On success the caller
// owns the lock and must release it (leaf_release) or commit (version + 2).
template <int F>
STM_DEV bool tx_lock_leaf(const TreeView& t, uint32_t id, WNode<F>& n) {
  if (id >= t.capacity) return false;
  NodeMeta* m = t.meta + id;
  uint64_t old = 0;
  if (lane_id() == 0)
    old = __hip_atomic_fetch_or(&m->lock, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  old = readlane64(old, 0);
  if (old & 1) return false;  // held by another transaction; the word is unchanged
  wave_order();               // content loads issue only after the lock is held
  const uint64_t info = uniform64(ld_relaxed(&m->info));
  n.low = uniform64(ld_relaxed(&m->low));
  n.high = uniform64(ld_relaxed(&m->high));
  n.next = (uint32_t)uniform64(ld_relaxed(&m->next));
  const uint64_t* kp = t.keys + (size_t)id * F;
  const uint64_t* vp = t.vals + (size_t)id * F;
#pragma unroll
  for (int j = 0; j < WNode<F>::SPL; ++j) {
    const int s = j * WAVE + lane_id();
    n.k[j] = s < F ? ld_relaxed(kp + s) : KEY_EMPTY;
    n.v[j] = s < F ? ld_relaxed(vp + s) : 0;
  }
  n.ver = old;
  n.id = id;
  n.count = meta_count(info);
  n.level = meta_level(info);
  return true;
}

// Gives a leaf lock back untouched (nothing was written under it).
STM_DEV void leaf_release(uint64_t* lw, uint64_t ver) {
  if (lane_id() == 0) st_relaxed(lw, ver);
}

// Upsert of (key, val) into leaf n, which this wavefront has locked (tx_lock_leaf).  pn
// is its parent when has_par (else n is the root).  A leaf's entries are unordered: a found key
// takes its value in place, a new one goes to slot count; a full leaf splits by rank, the lower
// half of the F + 1 entries (in key order) staying, the upper half moving to a new right
// sibling, each written in key order.  Every return other than TX_DONE and TX_NEED_PARENT has
// released the leaf; TX_NEED_PARENT keeps it locked.
template <int F>
STM_DEV TxResult tx_leaf_upsert(const TreeView& t, const WNode<F>& n, const WNode<F>& pn,
                                bool has_par, bool par_fresh, uint64_t hver, uint64_t key,
                                uint64_t val, bool& found, uint64_t& old, TxCounters& c) {
  constexpr int SPL = WNode<F>::SPL;
  const int lane = lane_id();
  const uint32_t s = wave_find<F>(n.k, key);
  found = s != NODE_NONE;
  old = found ? wave_at<F>(n.v, s) : 0;
  uint64_t* lw = &t.meta[n.id].lock;
  uint64_t* kp = t.keys + (size_t)n.id * F;
  uint64_t* vp = t.vals + (size_t)n.id * F;
  if (found) {
    if (lane == 0) st_relaxed(vp + s, val);
    tx_unlock(lw, n.ver, nullptr, 0, nullptr, 0);
    c.commits++;
    return TX_DONE;
  }
  if (n.count < (uint32_t)F) {
    if (lane == 0) {
      st_relaxed(kp + n.count, key);
      st_relaxed(vp + n.count, val);
    }
    if (lane == 1) st_relaxed(&t.meta[n.id].info, make_info(n.count + 1, 0));
    tx_unlock(lw, n.ver, nullptr, 0, nullptr, 0);
    c.commits++;
    c.new_keys++;
    return TX_DONE;
  }
  // Full leaf: split F+1 entries into h (left) and F+1-h (right).
  if (has_par && !par_fresh) return TX_NEED_PARENT;
  const uint32_t h = (F + 1) / 2;
  uint32_t ps = 0;
  if (has_par) {
    if (pn.count >= (uint32_t)F) {
      leaf_release(lw, n.ver);
      return TX_ABORT;
    }
    ps = parent_slot<F>(pn, key, n.id);
    if (ps == NODE_NONE) {
      leaf_release(lw, n.ver);
      return TX_ABORT;
    }
  }
  // Places in key order: an old entry's rank among the old ones, plus one above the new key;
  // the new key's place is the number of old entries below it.
  uint32_t r[SPL];
  wave_ranks<F>(n.k, F, r);
#pragma unroll
  for (int j = 0; j < SPL; ++j) r[j] += key < n.k[j] ? 1u : 0u;
  const uint32_t rn = wave_count_below<F>(n.k, F, key);
  // the separator: the entry of place h
  uint64_t cand = 0;
  bool has = false;
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    if (j * WAVE + lane < F && r[j] == h) {
      cand = n.k[j];
      has = true;
    }
  }
  const uint64_t sep = rn == h ? key : readlane64(cand, __builtin_ctzll(ballot(has) | (1ull << 63)));
  uint64_t* pw = has_par ? &t.meta[pn.id].lock : nullptr;
  uint64_t* hw = has_par ? nullptr : &t.hdr->lock;
  if (!tx_lock(pw, has_par ? pn.ver : 0, hw, hver, nullptr, 0)) {
    leaf_release(lw, n.ver);
    return TX_ABORT;
  }
  const uint32_t base = tx_alloc(t, has_par ? 1u : 2u);
  if (base == NODE_NONE) {
    tx_unlock(lw, n.ver - 2, pw, has_par ? pn.ver - 2 : 0, hw, hver - 2);
    return TX_FAIL;
  }
  const uint32_t rid = base;
  const uint32_t qid = has_par ? NODE_NONE : base + 1;
  tx_claim_new(t, rid, qid);
  uint64_t* rkp = t.keys + (size_t)rid * F;
  uint64_t* rvp = t.vals + (size_t)rid * F;
  // Every entry to its place: [0, h) in this leaf, [h, F+1) in the right sibling.  The left
  // leaf's slots from h up are freed; the right sibling's slots past its F+1-h entries are free.
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const uint32_t sl = (uint32_t)(j * WAVE + lane);
    if (sl < (uint32_t)F) {
      const uint32_t to = r[j];
      if (to >= h) {
        st_relaxed(rkp + (to - h), n.k[j]);
        st_relaxed(rvp + (to - h), n.v[j]);
      } else if (to != sl) {
        st_relaxed(kp + to, n.k[j]);
        st_relaxed(vp + to, n.v[j]);
      }
      if (sl >= h) st_relaxed(kp + sl, KEY_EMPTY);
      if (sl >= F + 1 - h) st_relaxed(rkp + sl, KEY_EMPTY);
    }
  }
  if (lane == 0) {
    if (rn >= h) {
      st_relaxed(rkp + (rn - h), key);
      st_relaxed(rvp + (rn - h), val);
    } else {
      st_relaxed(kp + rn, key);
      st_relaxed(vp + rn, val);
    }
  }
  tx_write_meta(t, rid, F + 1 - h, 0, sep, n.high, n.next);
  tx_write_meta(t, n.id, h, 0, n.low, sep, rid);
  if (has_par) {
    uint64_t pk[SPL], pv[SPL];
    parent_with_child<F>(pn, ps, sep, rid, pk, pv);
    tx_write_node<F>(t, pn.id, pk, pv, pn.count + 1, pn.level, pn.low, pn.high,
                     pn.next);
    tx_unlock(lw, n.ver, pw, pn.ver, &t.meta[rid].lock, 0ull);
  } else {
    uint64_t qk[SPL], qv[SPL];
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const int q = j * WAVE + lane;
      qk[j] = q == 0 ? n.low : (q == 1 ? sep : KEY_EMPTY);
      qv[j] = q == 0 ? (uint64_t)n.id : (q == 1 ? (uint64_t)rid : 0);
    }
    tx_write_node<F>(t, qid, qk, qv, 2, 1, n.low, n.high, NODE_NONE);
    if (lane == 0) {
      st_relaxed(&t.hdr->root, (uint64_t)qid);
      st_relaxed(&t.hdr->height, 2ull);
    }
    tx_unlock(lw, n.ver, &t.meta[rid].lock, 0ull, &t.hdr->lock, hver);
    if (lane == 0) st_relaxed(&t.meta[qid].lock, 2ull);
    c.root_splits++;
  }
  if (t.rank_of && lane == 0) {
    const uint32_t rr = t.rank_of[n.id];
    if (rr != NODE_NONE) t.dir[rr].pflags |= DIR_SPLIT;
  }
  c.leaf_splits++;
  c.commits++;
  c.new_keys++;
  c.split_happened = true;
  return TX_DONE;
}

// The inner node of level 1 that routed the wavefront's previous upsert.  Its content may
// be stale by the time it is used again; a stale route is caught by the leaf's fence check
// (and, for a split, by re-reading the parent with its version).
template <int F>
struct ParentCache {
  WNode<F> n;
  bool valid = false;
  bool fresh = false;  // read (version, then content) during the current attempt
};

// Routing part of an upsert attempt: the leaf that holds key, locked (tx_lock_leaf) into cur,
// with its parent par when has_par (read with its version during this attempt when
// par_fresh) and, after a descent from the root, the header version hver.  TX_DONE: the
// leaf is locked and its fences hold key.  Any other result holds no lock.
template <int F>
STM_DEV TxResult tx_route_leaf(const TreeView& t, uint64_t key, TxCounters& c, ParentCache<F>& pc,
                               uint32_t par_hint, WNode<F>& par, WNode<F>& cur, bool& has_par,
                               bool& par_fresh, uint64_t& hver) {
  has_par = false;
  par_fresh = false;
  hver = 0;
  uint32_t leaf_id;
  pc.fresh = false;
  if (!(pc.valid && key >= pc.n.low && key < pc.n.high) && par_hint < t.capacity) {
    // The level-1 node the lookup kernel routed this key through: one read instead of a
    // descent from the root.  Its fences and level check it; the leaf's fences check the route.
    pc.valid = tx_read<F>(t, par_hint, pc.n) && pc.n.level == 1 && key >= pc.n.low && key < pc.n.high;
    pc.fresh = pc.valid;
    c.node_reads++;
  }
  if (pc.valid && key >= pc.n.low && key < pc.n.high) {
    // Route through the cached level-1 node; the tree never loses a level, so no header
    // read is needed.  The leaf's fences validate the route.
    const uint32_t cle = wave_count_le<F>(pc.n.k, key);
    if (cle == 0 || cle > pc.n.count) {
      pc.valid = false;
      return TX_ABORT;
    }
    leaf_id = (uint32_t)wave_at<F>(pc.n.v, cle - 1);
    par = pc.n;
    has_par = true;
    par_fresh = pc.fresh;
  } else {
    hver = uniform64(ld_relaxed(&t.hdr->lock));
    if (hver & 1) return TX_ABORT;
    wave_order();
    const uint32_t root = (uint32_t)uniform64(ld_relaxed(&t.hdr->root));
    const uint32_t height = (uint32_t)uniform64(ld_relaxed(&t.hdr->height));
    if (height == 0 || height > 64) return TX_ABORT;
    leaf_id = root;
    if (height >= 2) {
      uint32_t id = root;
      for (uint32_t lvl = height - 1; lvl >= 1; --lvl) {
        c.node_reads++;
        if (!tx_read<F>(t, id, cur)) return TX_ABORT;
        if (cur.level != lvl || key < cur.low || key >= cur.high) return TX_ABORT;
        if (cur.count >= (uint32_t)F) {
          pc.valid = false;
          return tx_split_inner<F>(t, cur, par, has_par, hver, key, c);
        }
        const uint32_t cle = wave_count_le<F>(cur.k, key);
        if (cle == 0 || cle > cur.count) return TX_ABORT;
        const uint64_t child = wave_at<F>(cur.v, cle - 1);
        if (child >= t.capacity) return TX_ABORT;
        par = cur;
        has_par = true;
        id = (uint32_t)child;
      }
      pc.n = par;
      pc.valid = true;
      par_fresh = true;
      leaf_id = id;
    }
  }
  c.node_reads++;
  if (!tx_lock_leaf<F>(t, leaf_id, cur)) return TX_ABORT;
  if (cur.level != 0 || cur.count > (uint32_t)F || key < cur.low || key >= cur.high) {
    leaf_release(&t.meta[leaf_id].lock, cur.ver);
    pc.valid = false;
    return TX_ABORT;
  }
  return TX_DONE;
}

// Re-reads the parent of the locked leaf cur with its version before a split (the leaf stays
// locked); false releases the leaf.
template <int F>
STM_DEV bool tx_fresh_parent(const TreeView& t, uint64_t key, const WNode<F>& cur, WNode<F>& par,
                             TxCounters& c, ParentCache<F>& pc) {
  c.node_reads++;
  if (!tx_read<F>(t, par.id, par) || par.level != 1 || key < par.low || key >= par.high) {
    leaf_release(&t.meta[cur.id].lock, cur.ver);
    pc.valid = false;
    return false;
  }
  return true;
}

// One attempt of the upsert transaction for (key, val).
template <int F>
STM_DEV TxResult tx_upsert_attempt(const TreeView& t, uint64_t key, uint64_t val, bool& found,
                                   uint64_t& old, TxCounters& c, ParentCache<F>& pc,
                                   uint32_t par_hint) {
  WNode<F> par, cur;
  bool has_par, par_fresh;
  uint64_t hver;
  TxResult r = tx_route_leaf<F>(t, key, c, pc, par_hint, par, cur, has_par, par_fresh, hver);
  if (r != TX_DONE) return r;
  r = tx_leaf_upsert<F>(t, cur, par, has_par, par_fresh, hver, key, val, found, old, c);
  if (r == TX_NEED_PARENT) {
    // The leaf splits: the parent's current version and content are needed.
    if (!tx_fresh_parent<F>(t, key, cur, par, c, pc)) return TX_ABORT;
    r = tx_leaf_upsert<F>(t, cur, par, true, true, hver, key, val, found, old, c);
  }
  if (r != TX_DONE || c.split_happened) pc.valid = false;
  c.split_happened = false;
  return r;
}

// Merged upsert of the ops of lanes R into leaf n, which this wavefront has locked (each lane's
// key k_l and value v_l; keys are distinct): one transaction writes them all.  A leaf's entries
// are unordered: each op's key is looked up by equality over the leaf's slots, a key already
// present takes its new value in place, and the new keys are appended from slot count in lane
// order.  With T = count + new keys <= F the leaf is rewritten in place; with F < T <= 2F it
// splits by rank, the lower T/2 entries (in key order) staying, the rest moving to a new right
// sibling, each half written in key order, and one separator goes into the parent, as the
// single-op split does.  Returns as tx_leaf_upsert; found_l / old_l are each lane's answer.
template <int F>
STM_DEV TxResult tx_leaf_merge(const TreeView& t, const WNode<F>& n, const WNode<F>& pn,
                               bool has_par, bool par_fresh, uint64_t hver, uint64_t key0,
                               uint64_t R, uint64_t k_l, uint64_t v_l, bool& found_l,
                               uint64_t& old_l, TxCounters& c) {
  constexpr int SPL = WNode<F>::SPL;
  const int lane = lane_id();
  const bool mine = (R >> lane) & 1ull;
  uint64_t* lw = &t.meta[n.id].lock;
  // The ops' keys in the leaf, one op at a time (wave-uniform): uv = the values after the found
  // ops' updates, fmask = the found ops, each found lane's old value.
  uint64_t uv[SPL];
#pragma unroll
  for (int j = 0; j < SPL; ++j) uv[j] = n.v[j];
  uint64_t fmask = 0;
  uint64_t oldv = 0;
  for (uint64_t rest = R; rest; rest &= rest - 1) {
    const int i = __builtin_ctzll(rest);
    const uint64_t ki = readlane64(k_l, i);
    const uint32_t s = wave_find<F>(n.k, ki);
    if (s != NODE_NONE) {
      const uint64_t vi = readlane64(v_l, i);
      const uint64_t ov = wave_at<F>(n.v, s);
#pragma unroll
      for (int j = 0; j < SPL; ++j)
        if ((uint32_t)(j * WAVE + lane) == s) uv[j] = vi;
      fmask |= 1ull << i;
      if (lane == i) oldv = ov;
    }
  }
  const bool fnd = (fmask >> lane) & 1ull;
  const uint64_t newm = R & ~fmask;
  const uint32_t ins = (uint32_t)__popcll(newm);
  const uint32_t T = n.count + ins;
  found_l = fnd;
  old_l = fnd ? oldv : 0;
  uint64_t* kp = t.keys + (size_t)n.id * F;
  uint64_t* vp = t.vals + (size_t)n.id * F;
  if (T <= (uint32_t)F) {
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const uint32_t sl = (uint32_t)(j * WAVE + lane);
      if (sl < n.count && uv[j] != n.v[j]) st_relaxed(vp + sl, uv[j]);
    }
    if (mine && !fnd) {
      const uint32_t P = n.count + (uint32_t)__popcll(newm & lanemask_lt());
      st_relaxed(kp + P, k_l);
      st_relaxed(vp + P, v_l);
    }
    if (ins && lane == 1) st_relaxed(&t.meta[n.id].info, make_info(T, 0));
    tx_unlock(lw, n.ver, nullptr, 0, nullptr, 0);
    c.commits++;
    c.new_keys += ins;
    return TX_DONE;
  }
  // Split: entries at places [0, h) in key order stay, [h, T) move to a new right sibling.
  if (has_par && !par_fresh) return TX_NEED_PARENT;
  const uint32_t h = T / 2;
  uint32_t ps = 0;
  if (has_par) {
    if (pn.count >= (uint32_t)F) {
      leaf_release(lw, n.ver);
      return TX_ABORT;
    }
    ps = parent_slot<F>(pn, key0, n.id);
    if (ps == NODE_NONE) {
      leaf_release(lw, n.ver);
      return TX_ABORT;
    }
  }
  // Places in key order: for an old entry, the old entries below it plus the new keys below it;
  // for a new key, the same.
  uint32_t r[SPL];
#pragma unroll
  for (int j = 0; j < SPL; ++j) r[j] = 0;
  uint32_t rn = 0;  // per lane: entries below k_l
  for (uint32_t i = 0; i < n.count; ++i) {
    const uint64_t ki = wave_at<F>(n.k, i);
#pragma unroll
    for (int j = 0; j < SPL; ++j) r[j] += ki < n.k[j] ? 1u : 0u;
    rn += ki < k_l ? 1u : 0u;
  }
  for (uint64_t rest = newm; rest; rest &= rest - 1) {
    const uint64_t ki = readlane64(k_l, __builtin_ctzll(rest));
#pragma unroll
    for (int j = 0; j < SPL; ++j) r[j] += ki < n.k[j] ? 1u : 0u;
    rn += ki < k_l ? 1u : 0u;
  }
  // the separator: the entry whose place is h
  uint64_t cand = 0;
  bool has = false;
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const uint32_t sl = (uint32_t)(j * WAVE + lane);
    if (sl < n.count && r[j] == h) {
      cand = n.k[j];
      has = true;
    }
  }
  if (mine && !fnd && rn == h) {
    cand = k_l;
    has = true;
  }
  const uint64_t sep = readlane64(cand, __builtin_ctzll(ballot(has) | (1ull << 63)));
  uint64_t* pw = has_par ? &t.meta[pn.id].lock : nullptr;
  uint64_t* hw = has_par ? nullptr : &t.hdr->lock;
  if (!tx_lock(pw, has_par ? pn.ver : 0, hw, hver, nullptr, 0)) {
    leaf_release(lw, n.ver);
    return TX_ABORT;
  }
  const uint32_t base = tx_alloc(t, has_par ? 1u : 2u);
  if (base == NODE_NONE) {
    tx_unlock(lw, n.ver - 2, pw, has_par ? pn.ver - 2 : 0, hw, hver - 2);
    return TX_FAIL;
  }
  const uint32_t rid = base;
  const uint32_t qid = has_par ? NODE_NONE : base + 1;
  tx_claim_new(t, rid, qid);
  uint64_t* rkp = t.keys + (size_t)rid * F;
  uint64_t* rvp = t.vals + (size_t)rid * F;
  // Every entry to its place; the left leaf's old slots from h up are freed, and the right
  // sibling's slots past its T - h entries are free.
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const uint32_t sl = (uint32_t)(j * WAVE + lane);
    if (sl < n.count) {
      const uint32_t to = r[j];
      if (to >= h) {
        st_relaxed(rkp + (to - h), n.k[j]);
        st_relaxed(rvp + (to - h), uv[j]);
      } else if (to != sl || uv[j] != n.v[j]) {
        st_relaxed(kp + to, n.k[j]);
        st_relaxed(vp + to, uv[j]);
      }
      if (sl >= h) st_relaxed(kp + sl, KEY_EMPTY);  // the left node ends at h
    }
    if (sl >= T - h && sl < (uint32_t)F) st_relaxed(rkp + sl, KEY_EMPTY);
  }
  if (mine && !fnd) {
    if (rn >= h) {
      st_relaxed(rkp + (rn - h), k_l);
      st_relaxed(rvp + (rn - h), v_l);
    } else {
      st_relaxed(kp + rn, k_l);
      st_relaxed(vp + rn, v_l);
    }
  }
  tx_write_meta(t, rid, T - h, 0, sep, n.high, n.next);
  tx_write_meta(t, n.id, h, 0, n.low, sep, rid);
  if (has_par) {
    uint64_t pk[SPL], pv[SPL];
    parent_with_child<F>(pn, ps, sep, rid, pk, pv);
    tx_write_node<F>(t, pn.id, pk, pv, pn.count + 1, pn.level, pn.low, pn.high, pn.next);
    tx_unlock(lw, n.ver, pw, pn.ver, &t.meta[rid].lock, 0ull);
  } else {
    uint64_t qk[SPL], qv[SPL];
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const int sl = j * WAVE + lane;
      qk[j] = sl == 0 ? n.low : (sl == 1 ? sep : KEY_EMPTY);
      qv[j] = sl == 0 ? (uint64_t)n.id : (sl == 1 ? (uint64_t)rid : 0);
    }
    tx_write_node<F>(t, qid, qk, qv, 2, 1, n.low, n.high, NODE_NONE);
    if (lane == 0) {
      st_relaxed(&t.hdr->root, (uint64_t)qid);
      st_relaxed(&t.hdr->height, 2ull);
    }
    tx_unlock(lw, n.ver, &t.meta[rid].lock, 0ull, &t.hdr->lock, hver);
    if (lane == 0) st_relaxed(&t.meta[qid].lock, 2ull);
    c.root_splits++;
  }
  if (t.rank_of && lane == 0) {
    const uint32_t r = t.rank_of[n.id];
    if (r != NODE_NONE) t.dir[r].pflags |= DIR_SPLIT;
  }
  c.leaf_splits++;
  c.commits++;
  c.new_keys += ins;
  c.split_happened = true;
  return TX_DONE;
}

// One attempt of a merged upsert: routes key0 (the lowest remaining lane's key) to its leaf,
// locks it and merges every remaining op of the round (todo) whose key the leaf's fences
// hold, at most F of them.  On TX_DONE, done is the lanes whose ops committed.  With a leaf
// hint (the leaf the lookup kernel found) and its level-1 hint the leaf is locked directly:
// fast-path commits never split, so the leaf still holds the key unless a transaction of this
// launch split it, which its fences show; a split then reads the hinted parent with its version
// and checks that it points at the leaf.
template <int F>
STM_DEV TxResult tx_merge_attempt(const TreeView& t, uint64_t key0, uint64_t todo, uint64_t k_l,
                                  uint64_t v_l, TxCounters& c, ParentCache<F>& pc,
                                  uint32_t par_hint, uint32_t leaf_hint, uint64_t& done,
                                  bool& found_l, uint64_t& old_l) {
  WNode<F> par, cur;
  bool has_par = false, par_fresh = false, direct = false;
  uint64_t hver = 0;
  TxResult r;
  if (leaf_hint < t.capacity && par_hint < t.capacity) {
    c.node_reads++;
    if (!tx_lock_leaf<F>(t, leaf_hint, cur)) return TX_ABORT;  // locked by another transaction
    if (cur.level == 0 && cur.count <= (uint32_t)F && key0 >= cur.low && key0 < cur.high) {
      direct = true;
      has_par = true;
      par.id = par_hint;
    } else {
      leaf_release(&t.meta[leaf_hint].lock, cur.ver);
    }
  }
  if (!direct) {
    r = tx_route_leaf<F>(t, key0, c, pc, par_hint, par, cur, has_par, par_fresh, hver);
    if (r != TX_DONE) return r;
  }
  uint64_t R = ballot(((todo >> lane_id()) & 1ull) && k_l >= cur.low && k_l < cur.high);
  if constexpr (F < WAVE) {
    // at most F ops, so that T <= 2F (the lowest lanes: key0's lane is the lowest of todo)
    uint64_t keep = 0;
    for (int q = 0; q < F && R; ++q) {
      keep |= R & (~R + 1);
      R &= R - 1;
    }
    R = keep;
  }
  r = tx_leaf_merge<F>(t, cur, par, has_par, par_fresh, hver, key0, R, k_l, v_l, found_l, old_l, c);
  if (r == TX_NEED_PARENT) {
    if (!tx_fresh_parent<F>(t, key0, cur, par, c, pc)) return TX_ABORT;
    r = tx_leaf_merge<F>(t, cur, par, true, true, hver, key0, R, k_l, v_l, found_l, old_l, c);
  }
  if (r != TX_DONE || c.split_happened) pc.valid = false;
  c.split_happened = false;
  done = r == TX_DONE ? R : 0;
  return r;
}

STM_DEV void backoff(uint32_t attempt) {
  const uint32_t spins = min(attempt, 32u);
  for (uint32_t i = 0; i < spins; ++i) __builtin_amdgcn_s_sleep(4);
}

// Runs one merged upsert to completion (or gives up after max_attempts) and answers the
// np ops of its group that wait for the tree (sorted positions gs .. gs+np-1).
template <int F>
STM_DEV bool run_upsert(const TreeView& t, uint64_t key, uint64_t val, uint32_t gs, uint32_t np,
                        const uint32_t* __restrict__ sidx, uint64_t* __restrict__ out_vals,
                        uint8_t* __restrict__ out_found, uint32_t max_attempts, TxCounters& c,
                        ParentCache<F>& pc, uint32_t par_hint = NODE_NONE) {
  bool found = false;
  uint64_t old = 0;
  uint32_t aborts = 0, restarts = 0;
  for (;;) {
    // The hint is only tried on the first attempt; retries descend from the root.
    const TxResult r = tx_upsert_attempt<F>(t, key, val, found, old, c, pc,
                                            aborts + restarts == 0 ? par_hint : NODE_NONE);
    if (r == TX_DONE) break;
    if (r == TX_FAIL) return true;  // pool exhausted: error bit raised, op dropped
    if (r == TX_ABORT) {
      c.aborts++;
      if (++aborts >= max_attempts) return false;
      backoff(aborts);
    } else if (++restarts >= 4 * max_attempts) {  // TX_RESTART after a preemptive split
      return false;
    }
  }
  if (out_vals) {
    for (uint32_t p = lane_id(); p < np; p += WAVE) {
      const uint32_t pos = pos_of(sidx[gs + p]);
      out_vals[pos] = old;
      out_found[pos] = found ? 1 : 0;
    }
  }
  return true;
}


STM_DEV void flush_counters(const TreeView& t, const TxCounters& c) {
  if (lane_id() == 0) {
    const uint32_t w = wave_stripe();
    stat_add(t.stats, ST_COMMITS, w, c.commits);
    stat_add(t.stats, ST_ABORTS, w, c.aborts);
    stat_add(t.stats, ST_LEAF_SPLITS, w, c.leaf_splits);
    stat_add(t.stats, ST_INNER_SPLITS, w, c.inner_splits);
    stat_add(t.stats, ST_ROOT_SPLITS, w, c.root_splits);
    stat_add(t.stats, ST_KEYS, w, c.new_keys);
    stat_add(t.stats, ST_NODE_READS, w, c.node_reads);
  }
}

// Also adds the wavefront's leaf splits to the batch's counter (the host's retrain trigger).
STM_DEV void flush_counters(const TreeView& t, const TxCounters& c, BatchCounters* cnt) {
  if (lane_id() == 0 && c.leaf_splits) atomicAdd(&cnt->leaf_splits, (uint32_t)c.leaf_splits);
  flush_counters(t, c);
}

// Lane-group fast path: NG = 64 / G merged INSERTs of one wavefront at once, each by G lanes.
// It is the leaf-only transaction of the wave-wide path (update in place, or append at the
// leaf's count: a leaf's entries are unordered), commit with version + 2, run for NG ops in
// lock-step, started from the leaf, slot and entry count the lookup kernel found for the key,
// and it runs only for the op that owns its leaf in this launch: the first merged INSERT of the
// leaf in key order.  Groups are sorted by key, so an op owns its leaf when none of the groups
// between the leaf's low fence and its own key is an INSERT; lane g of the op's lane group reads
// group gi - 1 - g (key and merged op), and ownership is proven when one of those G keys lies
// below the fence (or the batch starts).  No other group of the launch then writes the leaf, and
// no writer ran since the lookup kernel read it, so the hint is exact and no lock is taken.  An
// op that does not own its leaf (another INSERT of the batch precedes it there, or more than G
// groups of the leaf do), whose leaf is full or whose hint disagrees with the leaf's version,
// count or fences is handed back to the wave-wide path, which runs after this launch.  A found
// key reads and writes its one value; a new key writes one key and one value.  (Measured with
// sorted leaves, whose inserts shifted the slots above the key: 97 µs per 2^22-op batch, ≈56 of
// them the shifted slots' stores; with the leaf's test-and-set instead of the ownership test
// before that, 131 µs.)
// src = round lane of this group's op (-1: none).  Returns the round lanes whose op committed.
template <int F, int G>
STM_DEV uint64_t fast_leaf_upserts(const TreeView& t, int src, uint64_t k_l, uint64_t v_l,
                                   uint32_t leaf_l, uint64_t ver_l, uint32_t up_l, uint32_t gi_l,
                                   const uint64_t* __restrict__ ukey,
                                   const uint8_t* __restrict__ uop, uint32_t gs_l,
                                   uint32_t np_l, uint32_t p0_l, const uint32_t* __restrict__ sidx,
                                   uint64_t* __restrict__ out_vals,
                                   uint8_t* __restrict__ out_found, TxCounters& c) {
  const int lane = lane_id();
  const int g = lane & (G - 1);
  const int gbase = lane - g;
  const int sl = src >= 0 ? src : 0;
  const uint64_t key = shfl64(k_l, sl);
  const uint64_t val = shfl64(v_l, sl);
  const uint32_t leaf = (uint32_t)__shfl((int)leaf_l, sl, WAVE);
  const uint32_t gs = (uint32_t)__shfl((int)gs_l, sl, WAVE);
  const uint32_t np = (uint32_t)__shfl((int)np_l, sl, WAVE);
  const uint32_t p0 = (uint32_t)__shfl((int)p0_l, sl, WAVE);
  const uint32_t up = (uint32_t)__shfl((int)up_l, sl, WAVE);
  const uint32_t gi = (uint32_t)__shfl((int)gi_l, sl, WAVE);
  const uint64_t ver = shfl64(ver_l, sl);
  const uint32_t pos = up & UPOS_SLOT;  // the key's slot when found
  const bool hfound = (up & UPOS_FOUND) != 0;
  const uint32_t hcount = up >> UPOS_COUNT_SHIFT;
  // A hint out of range (never written by the lookup kernel) is not acted on.
  const bool act = src >= 0 && leaf < t.capacity && hcount <= (uint32_t)F &&
                   (!hfound || pos < hcount);
  const size_t base = (size_t)(act ? leaf : 0u) * F;
  NodeMeta* m = t.meta + (act ? leaf : 0u);
  uint64_t info = 0, low = 0, high = 0, word = 0, old = 0;
  // The leaf's metadata line, the found key's value and the G groups before the op go out
  // together.
  const bool has_prev = act && (uint32_t)g < gi;
  uint64_t pkey = 0;
  uint8_t pop = UOP_SKIP;
  if (act) {
    word = m->lock;
    info = m->info;
    low = m->low;
    high = m->high;
    if (has_prev) {
      pkey = ukey[gi - 1 - (uint32_t)g];
      pop = uop[gi - 1 - (uint32_t)g];
    }
    if (hfound && g == 0) old = t.vals[base + pos];
  }
  old = shfl64(old, gbase);
  // Keys fall as g grows, so the groups inside the leaf's range are a prefix of the G read.
  const bool in_leaf = has_prev && pkey >= low;
  const uint64_t gmask = (G == WAVE ? ~0ull : ((1ull << G) - 1)) << gbase;
  const uint64_t inside = ballot(in_leaf) & gmask;
  const bool owner = (ballot(in_leaf && pop == OP_INSERT) & gmask) == 0 && inside != gmask;
  const bool held = act && owner && !(word & 1ull);
  const bool fresh = held && word == ver;
  const uint32_t count = meta_count(info);
  const bool fits = fresh && meta_level(info) == 0 && count == hcount && key >= low && key < high;
  const bool found = fits && hfound;
  const bool insert = fits && !hfound && count < (uint32_t)F;
  const bool commit = found || insert;
  if (found) {
    if (g == 0) t.vals[base + pos] = val;
  } else {
    old = 0;
  }
  if (insert && g == 0) {  // slot count is free (KEY_EMPTY)
    t.keys[base + count] = key;
    t.vals[base + count] = val;
  }
  if (insert && g == 1 % G) st_relaxed(&m->info, make_info(count + 1, 0));
  if (commit && g == 0) st_relaxed(&m->lock, word + 2);
  if (commit) {
    for (uint32_t p = (uint32_t)g; p < np; p += G) {
      const uint32_t pos_o = p == 0 ? p0 : pos_of(sidx[gs + p]);
      out_vals[pos_o] = old;
      out_found[pos_o] = found ? 1 : 0;
    }
  }
  const uint64_t leaders = ballot(commit && g == 0);
  const uint64_t news = ballot(insert && g == 0);
  c.commits += (uint32_t)__popcll(leaders);
  c.new_keys += (uint32_t)__popcll(news);
  c.node_reads += (uint32_t)__popcll(ballot(act && g == 0));
  uint64_t done = 0;
#pragma unroll
  for (int q = 0; q < WAVE / G; ++q) {
    const int sq = __builtin_amdgcn_readlane(src, q * G);
    if ((leaders >> (q * G)) & 1ull) done |= 1ull << sq;
  }
  return done;
}

// Fast-path kernel over the batch's merged INSERTs (ilist, n_upserts entries): each wavefront
// owns a contiguous part and runs it NG ops at a time on the lane-group fast path.  The ops it
// hands back go to the same part of slist, from its start, and their number to scount[wave]:
// no counter is shared (one hot address took an atomic from every wavefront at once, at the
// end of the launch).
// Answers of the hot GET groups the lookup kernel queued (more than FANOUT_DIRECT waiting
// ops): chunk c of the queue is op range [first, first + FANOUT_CHUNK) of group g's waiting
// ops, written by one workgroup, chunks spread over the grid.  Runs at the start of the upsert
// fast path, whose writes (INSERT groups' answers) never touch a GET group's ops.  Reference:
// the per-op answer write-back, writebackAns (src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:46-56).
STM_DEV void fanout_answers(const uint64_t* __restrict__ fanout, uint64_t cap,
                            const BatchCounters* cnt, const uint32_t* __restrict__ gstart,
                            const uint32_t* __restrict__ npend, const uint64_t* __restrict__ gval,
                            const uint32_t* __restrict__ sidx, uint64_t* __restrict__ out_vals,
                            uint8_t* __restrict__ out_found) {
  const uint32_t nc = (uint32_t)min<uint64_t>(cnt->n_fanout, cap);
  for (uint32_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const uint64_t e = fanout[c];
    const uint32_t g = (uint32_t)e;
    const uint32_t first = (uint32_t)(e >> 32) & 0x7fffffffu;
    const uint8_t found = (uint8_t)(e >> 63);
    const uint32_t gs = gstart[g];
    const uint32_t end = min(npend[g], first + FANOUT_CHUNK);
    const uint64_t v = gval[g];
    for (uint32_t p = first + threadIdx.x; p < end; p += blockDim.x) {
      const uint32_t pos = pos_of(sidx[gs + p]);
      out_vals[pos] = v;
      out_found[pos] = found;
    }
  }
}

template <int F, int G>
__global__ __launch_bounds__(256) void k_upsert_fast(TreeView t, const uint64_t* __restrict__ ukey,
                                                     const uint64_t* __restrict__ uval,
                                                     const uint8_t* __restrict__ uop,
                                                     const uint32_t* __restrict__ gstart,
                                                     const uint32_t* __restrict__ npend,
                                                     const uint32_t* __restrict__ sidx,
                                                     const uint32_t* __restrict__ ilist,
                                                     const InsHint* __restrict__ ihint,
                                                     BatchCounters* cnt, uint32_t* __restrict__ slist,
                                                     uint32_t* __restrict__ scount,
                                                     uint64_t* __restrict__ out_vals,
                                                     uint8_t* __restrict__ out_found,
                                                     const BatchDesc* __restrict__ desc,
                                                     const uint64_t* __restrict__ fanout,
                                                     uint64_t fanout_cap,
                                                     const uint64_t* __restrict__ gval) {
  constexpr int NG = WAVE / G;
  const uint32_t N = cnt->n_upserts;
  if (desc) {
    out_vals = desc->out_vals;
    out_found = desc->out_found;
  }
  if (fanout) fanout_answers(fanout, fanout_cap, cnt, gstart, npend, gval, sidx, out_vals, out_found);
  const uint32_t nw = gridDim.x * (blockDim.x / WAVE);
  const uint32_t w = uniform32((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
  const uint32_t per = (N + nw - 1) / nw;
  const uint32_t e0 = min(w * per, N);
  const uint32_t e1 = min(e0 + per, N);
  TxCounters c;
  uint32_t upserts = 0;
  uint32_t handed = 0;  // ops handed back so far (slist[e0 .. e0 + handed))
  for (uint32_t base = e0; base < e1; base += WAVE) {
    const uint32_t e = base + lane_id();
    const bool mine = e < e1;
    // Unconditional loads from an index clamped into the range (e1 > base), selected
    // afterwards, so the gathers issue together.
    const uint32_t gi = ilist[mine ? e : e1 - 1];
    uint64_t todo = ballot(mine);
    const InsHint h = ihint[gi];
    const uint64_t v_raw = uval[gi];
    const uint64_t k_raw = ukey[gi];
    const uint32_t gs_raw = h.gs;
    const uint32_t np_raw = h.np;
    const uint32_t leaf_raw = h.leaf;
    const uint64_t ver_raw = h.ver;
    const uint32_t up_raw = h.pos;
    const uint64_t k_l = mine ? k_raw : 0;
    const uint64_t v_l = mine ? v_raw : 0;
    const uint32_t gs_l = mine ? gs_raw : 0;
    const uint32_t np_l = mine ? np_raw : 0;
    const uint32_t leaf_l = mine ? leaf_raw : NODE_NONE;
    const uint64_t ver_l = mine ? ver_raw : 1ull;
    const uint32_t up_l = mine ? up_raw : 0u;
    // The first answered op's input position, loaded with the op data rather than inside
    // the round (gs_l is a sorted position of the batch).
    const uint32_t sx_raw = sidx[gs_l];
    const uint32_t p0_l = (mine && np_l) ? pos_of(sx_raw) : 0u;
    for (uint64_t rest = todo; rest;) {
      // group q takes the q-th remaining op of the round
      int my_src = -1;
#pragma unroll
      for (int q = 0; q < NG; ++q) {
        const int l = rest ? __builtin_ctzll(rest) : -1;
        if (l >= 0) rest &= rest - 1;
        if (lane_id() / G == q) my_src = l;
      }
      const uint64_t done = fast_leaf_upserts<F, G>(t, my_src, k_l, v_l, leaf_l, ver_l, up_l, gi,
                                                    ukey, uop, gs_l, np_l, p0_l, sidx, out_vals,
                                                    out_found, c);
      upserts += (uint32_t)__popcll(done);
      todo &= ~done;
    }
    if ((todo >> lane_id()) & 1ull)  // handed back
      slist[e0 + handed + (uint32_t)__popcll(todo & lanemask_lt())] = gi;
    handed += (uint32_t)__popcll(todo);
  }
  if (lane_id() == 0) scount[w] = handed;
  flush_counters(t, c);
  if (lane_id() == 0) stat_add(t.stats, ST_UPSERTS, wave_stripe(), upserts);
}

// Answers of the committed ops of a round (lanes with ans): a lane writes its own group's few
// waiting ops (np of them from sorted position gs); a group with many (a hot key) is written by
// the whole wavefront.  Called by every lane of the wavefront.
STM_DEV void write_upsert_answers(bool ans, bool found, uint64_t old, uint32_t gs, uint32_t np,
                                  const uint32_t* __restrict__ sidx,
                                  uint64_t* __restrict__ out_vals, uint8_t* __restrict__ out_found) {
  const bool many = ans && np > 8;
  if (ans && !many) {
    for (uint32_t p = 0; p < np; ++p) {
      const uint32_t pos = pos_of(sidx[gs + p]);
      out_vals[pos] = old;
      out_found[pos] = found ? 1 : 0;
    }
  }
  for (uint64_t bm = ballot(many); bm; bm &= bm - 1) {
    const int b = __builtin_ctzll(bm);
    const uint32_t gsb = __builtin_amdgcn_readlane(gs, b);
    const uint32_t npb = __builtin_amdgcn_readlane(np, b);
    const uint64_t ob = readlane64(old, b);
    const uint8_t fb = (uint8_t)__builtin_amdgcn_readlane(found ? 1 : 0, b);
    for (uint32_t p = lane_id(); p < npb; p += WAVE) {
      const uint32_t pos = pos_of(sidx[gsb + p]);
      out_vals[pos] = ob;
      out_found[pos] = fb;
    }
  }
}

// Wave-wide transactions for the ops the fast path handed back, or for every merged INSERT
// of the batch when slist is null.  The handed-back ops lie in the insert list's index range
// [0, n_upserts): fast-path wavefront f owned [f * per_f, (f + 1) * per_f) and wrote its
// scount[f] ops from the start of that range.  Each wavefront here owns a contiguous part of
// the index range (key order within a part), so consecutive upserts mostly reuse the cached
// parent.
template <int F>
__global__ __launch_bounds__(256) void k_upsert(TreeView t, const uint64_t* __restrict__ ukey,
                                                const uint64_t* __restrict__ uval,
                                                const uint8_t* __restrict__ uop,
                                                const uint32_t* __restrict__ gstart,
                                                const uint32_t* __restrict__ npend,
                                                const uint32_t* __restrict__ sidx,
                                                const uint32_t* __restrict__ slist,
                                                const uint32_t* __restrict__ scount,
                                                uint32_t fast_waves,
                                                const InsHint* __restrict__ ihint,
                                                BatchCounters* cnt, uint32_t* __restrict__ dlist,
                                                uint32_t max_attempts,
                                                uint64_t* __restrict__ out_vals,
                                                uint8_t* __restrict__ out_found,
                                                const BatchDesc* __restrict__ desc) {
  const uint32_t U = slist ? cnt->n_upserts : cnt->n_unique;
  if (U == 0) return;
  const uint32_t per_f = slist ? (U + fast_waves - 1) / fast_waves : 1u;
  if (desc) {
    out_vals = desc->out_vals;
    out_found = desc->out_found;
  }
  const uint32_t nw = gridDim.x * (blockDim.x / WAVE);
  const uint32_t w = uniform32((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
  TxCounters c;
  ParentCache<F> pc;
  uint32_t upserts = 0, deferred = 0;
  // The wavefront's share, as ranges [s0, s1) of the group index (no slist) or of slist.
  // Without slist: an equal contiguous part of the groups.  With slist: fast-path wavefront f
  // handed back scount[f] ops at slist[f * per_f ...]; with r = nw / fast_waves wavefronts per
  // such range each takes a contiguous r-th of it (so the hand-backs, which sit at the start of
  // each range, spread over all wavefronts), with fewer wavefronts than ranges wavefront w takes
  // ranges w, w + nw, ...
  for (uint32_t f = w;; f += nw) {
    uint32_t s0, s1;
    if (!slist) {
      if (f != w) break;
      const uint32_t per = (U + nw - 1) / nw;
      s0 = min(w * per, U);
      s1 = min(s0 + per, U);
    } else if (nw >= fast_waves) {
      if (f != w) break;
      const uint32_t r = nw / fast_waves;
      const uint32_t fi = w / r, q = w % r;
      if (fi >= fast_waves) break;
      const uint32_t n_f = uniform32(scount[fi]);
      const uint32_t chunk = (n_f + r - 1) / r;
      s0 = fi * per_f + min(q * chunk, n_f);
      s1 = fi * per_f + min((q + 1) * chunk, n_f);
    } else {
      if (f >= fast_waves) break;
      s0 = f * per_f;
      s1 = s0 + uniform32(scount[f]);
    }
  for (uint32_t base = s0; base < s1; base += WAVE) {
    // Op data of the round's upserts, loaded together and read lane by lane.
    const uint32_t e = base + lane_id();
    const bool in = e < s1;
    const uint32_t gi = in ? (slist ? slist[e] : e) : 0;
    const bool mine = in && (slist || uop[gi] == OP_INSERT);
    uint64_t todo = ballot(mine);
    // The op: from the lookup kernel's record when there is one, else from the fold's arrays
    // (then without leaf hints).
    uint64_t k_l = 0, v_l = 0;
    uint32_t gs_l = 0, np_l = 0, par_l = NODE_NONE, leaf_l = NODE_NONE, up_l = 0;
    if (ihint) {
      const InsHint h = ihint[gi];
      const uint64_t k = ukey[gi];
      const uint64_t v = uval[gi];
      if (mine) {
        k_l = k;
        v_l = v;
        gs_l = h.gs;
        np_l = h.np;
        par_l = h.par;
        leaf_l = h.leaf;
        up_l = h.pos;
      }
    } else if (mine) {
      k_l = ukey[gi];
      v_l = uval[gi];
      gs_l = gstart[gi];
      np_l = npend[gi];
    }
    if (ihint) {
      // Every lane at once: its own leaf transaction on the hinted leaf, for an op that needs no
      // split (leaf entries are unordered, so a new key is one append and a found key one value
      // store).  Test-and-set of the leaf's lock word; under the lock the leaf must still be a
      // leaf holding the key in its fences (a split by a transaction of this launch moves keys
      // out), a found key must still sit at its hinted slot (a split rewrites the slots in key
      // order) and a new key needs a free slot.  The same lock word serializes these with the
      // wave-wide transactions below, which take every op that fails here.
      bool held = false, ok = false, fnd = false;
      uint64_t w = 0, info = 0, low = 0, high = 0, kat = 0, old = 0;
      const uint32_t pos = up_l & UPOS_SLOT;
      const bool hfound = (up_l & UPOS_FOUND) != 0;
      NodeMeta* m = t.meta + (leaf_l < t.capacity ? leaf_l : 0u);
      const size_t nb = (size_t)(leaf_l < t.capacity ? leaf_l : 0u) * F;
      if (mine && leaf_l < t.capacity && (!hfound || pos < (uint32_t)F)) {
        w = __hip_atomic_fetch_or(&m->lock, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        held = !(w & 1ull);
      }
      wave_order();  // the leaf's content is read only after its lock is held
      if (held) {
        info = ld_relaxed(&m->info);
        low = ld_relaxed(&m->low);
        high = ld_relaxed(&m->high);
        if (hfound) {
          kat = ld_relaxed(t.keys + nb + pos);
          old = ld_relaxed(t.vals + nb + pos);
        }
      }
      const uint32_t count = meta_count(info);
      if (held && meta_level(info) == 0 && count <= (uint32_t)F && k_l >= low && k_l < high) {
        if (hfound) {
          ok = kat == k_l;
          fnd = ok;
          if (ok) st_relaxed(t.vals + nb + pos, v_l);
        } else if (count < (uint32_t)F) {
          ok = true;
          st_relaxed(t.keys + nb + count, k_l);
          st_relaxed(t.vals + nb + count, v_l);
          st_relaxed(&m->info, make_info(count + 1, 0));
        }
      }
      wave_order();  // the content stores are acknowledged before the lock word changes
      if (held) st_relaxed(&m->lock, ok ? w + 2 : w);
      const uint64_t adone = ballot(ok);
      write_upsert_answers(ok, fnd, fnd ? old : 0ull, gs_l, np_l, sidx, out_vals, out_found);
      c.commits += (uint32_t)__popcll(adone);
      c.new_keys += (uint32_t)__popcll(ballot(ok && !fnd));
      c.node_reads += (uint32_t)__popcll(ballot(held));
      upserts += (uint32_t)__popcll(adone);
      todo &= ~adone;
    }
    // Ops of the round whose keys share a leaf commit together (tx_merge_attempt): a hot
    // leaf takes one transaction per round instead of one per op.
    while (todo) {
      const int l = __builtin_ctzll(todo);
      const uint64_t key = readlane64(k_l, l);
      const uint32_t ph = __builtin_amdgcn_readlane(par_l, l);
      const uint32_t lh = __builtin_amdgcn_readlane(leaf_l, l);
      uint32_t aborts = 0, restarts = 0;
      uint64_t done = 0;
      bool found_l = false;
      uint64_t old_l = 0;
      bool gave_up = false;
      for (;;) {
        // The hint is only tried on the first attempt; retries descend from the root.
        const bool first = aborts + restarts == 0;
        const TxResult r = tx_merge_attempt<F>(t, key, todo, k_l, v_l, c, pc,
                                               first ? ph : NODE_NONE, first ? lh : NODE_NONE,
                                               done, found_l, old_l);
        if (r == TX_DONE) break;
        if (r == TX_FAIL) {  // pool exhausted: error bit raised, the op is dropped
          done = 1ull << l;
          found_l = false;
          old_l = 0;
          break;
        }
        if (r == TX_ABORT) {
          c.aborts++;
          if (++aborts >= max_attempts) {
            gave_up = true;
            break;
          }
          backoff(aborts);
        } else if (++restarts >= 4 * max_attempts) {  // TX_RESTART after a preemptive split
          gave_up = true;
          break;
        }
      }
      if (gave_up) {  // the op of lane l goes to the serial pass
        uint32_t slot = 0;
        if (lane_id() == 0) slot = atomicAdd(&cnt->n_deferred, 1u);
        slot = uniform32(__shfl(slot, 0, WAVE));
        const uint32_t g = __builtin_amdgcn_readlane(gi, l);
        if (lane_id() == 0) dlist[slot] = g;
        deferred++;
        upserts++;
        todo &= ~(1ull << l);
        continue;
      }
      write_upsert_answers(((done >> lane_id()) & 1ull) != 0, found_l, old_l, gs_l, np_l, sidx,
                           out_vals, out_found);
      upserts += (uint32_t)__popcll(done);
      todo &= ~done;
    }
  }
  }  // ranges
  flush_counters(t, c, cnt);
  if (lane_id() == 0) {
    stat_add(t.stats, ST_UPSERTS, wave_stripe(), upserts);
    stat_add(t.stats, ST_DEFERRED, wave_stripe(), deferred);
  }
}

// Serial pass over the ops that exhausted their attempts: one wavefront, no competitors,
// so every transaction commits on its first attempt.
template <int F>
__global__ __launch_bounds__(64) void k_upsert_serial(TreeView t, const uint64_t* __restrict__ ukey,
                                                      const uint64_t* __restrict__ uval,
                                                      const uint32_t* __restrict__ gstart,
                                                      const uint32_t* __restrict__ npend,
                                                      const uint32_t* __restrict__ sidx,
                                                      const uint32_t* __restrict__ dlist,
                                                      const BatchCounters* cnt,
                                                      uint64_t* __restrict__ out_vals,
                                                      uint8_t* __restrict__ out_found,
                                                      const BatchDesc* __restrict__ desc,
                                                      BatchCounters* host_out) {
  const uint32_t n = cnt->n_deferred;
  if (desc) {
    out_vals = desc->out_vals;
    out_found = desc->out_found;
  }
  TxCounters c;
  ParentCache<F> pc;
  for (uint32_t e = 0; e < n; ++e) {
    const uint32_t g = dlist[e];
    if (!run_upsert<F>(t, ukey[g], uval[g], gstart[g], npend[g], sidx, out_vals, out_found, 1024u,
                       c, pc)) {
      if (lane_id() == 0) atomicOr((unsigned long long*)&t.hdr->error, (unsigned long long)ERR_LIVELOCK);
    }
  }
  flush_counters(t, c, const_cast<BatchCounters*>(cnt));
  // The last kernel of a batch: the tree's error bits travel back with the batch counters.
  wave_order();
  const uint64_t err = ld_relaxed(&t.hdr->error);
  if (lane_id() == 0) const_cast<BatchCounters*>(cnt)->hdr_error = err;
  if (host_out) {
    // Lanes 0-7 store the 64-byte block into pinned host memory, which spares the batch a
    // separate copy on the stream.  Agent-scope loads read the other kernels' counter updates
    // past this CU's cache; the system fence makes the stores visible before the kernel ends.
    static_assert(sizeof(BatchCounters) == 64 && offsetof(BatchCounters, hdr_error) == 40,
                  "hdr_error is word 5");
    const int l = lane_id();
    if (l < 8) {
      const uint64_t w = l == 5 ? err : ld_relaxed(reinterpret_cast<const uint64_t*>(cnt) + l);
      __hip_atomic_store(reinterpret_cast<uint64_t*>(host_out) + l, w, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __threadfence_system();
  }
}

// ================================================================== dump / check
__global__ void k_collect_leaves(TreeView t, uint32_t n_nodes, uint64_t* __restrict__ lows,
                                 uint32_t* __restrict__ ids, uint32_t* __restrict__ counter) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n_nodes) return;
  const NodeMeta m = t.meta[id];
  if (meta_level(m.info) != 0) return;
  const uint32_t slot = atomicAdd(counter, 1u);
  lows[slot] = m.low;
  ids[slot] = id;
}

__global__ void k_leaf_counts(TreeView t, const uint32_t* __restrict__ order,
                              const uint32_t* __restrict__ ids_unsorted, uint32_t n_leaves,
                              uint32_t* __restrict__ sorted_ids, uint32_t* __restrict__ counts) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_leaves) return;
  const uint32_t id = ids_unsorted[order[r]];
  sorted_ids[r] = id;
  counts[r] = meta_count(t.meta[id].info);
}

template <int F>
__global__ void k_dump_copy(TreeView t, const uint32_t* __restrict__ sorted_ids,
                            const uint64_t* __restrict__ offsets, uint32_t n_leaves,
                            uint64_t* __restrict__ out_k, uint64_t* __restrict__ out_v) {
  const uint64_t gi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gi >= (uint64_t)n_leaves * F) return;
  const uint32_t r = (uint32_t)(gi / F);
  const uint32_t s = (uint32_t)(gi % F);
  const uint64_t o0 = offsets[r];
  const uint64_t o1 = offsets[r + 1];
  if (o0 + s >= o1) return;
  const uint32_t id = sorted_ids[r];
  // A leaf's entries are unordered: slot s goes to its rank among the leaf's count entries.
  const uint64_t* kp = t.keys + (size_t)id * F;
  const uint64_t k = kp[s];
  const uint32_t count = (uint32_t)(o1 - o0);
  uint32_t rank = 0;
  for (uint32_t u = 0; u < count; ++u) rank += kp[u] < k ? 1u : 0u;
  out_k[o0 + rank] = k;
  out_v[o0 + rank] = t.vals[(size_t)id * F + s];
}

// One wavefront per node; counts violated invariants into *bad.
template <int F>
__global__ __launch_bounds__(256) void k_check(TreeView t, uint32_t n_nodes,
                                               unsigned long long* bad) {
  constexpr int SPL = WNode<F>::SPL;
  const uint32_t id = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (id >= n_nodes) return;
  const NodeMeta m = t.meta[id];
  const uint32_t count = meta_count(m.info);
  const uint32_t level = meta_level(m.info);
  uint32_t errs = 0;
  if (m.lock & 1) errs++;
  if (count > (uint32_t)F) {
    if (lane_id() == 0) atomicAdd(bad, 1ull);
    return;
  }
  if (m.low >= m.high) errs++;
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const int s = j * WAVE + lane_id();
    if (s >= F) continue;
    const uint64_t k = t.keys[(size_t)id * F + s];
    const uint64_t v = t.vals[(size_t)id * F + s];
    if ((uint32_t)s >= count) {
      if (k != KEY_EMPTY) errs++;
      continue;
    }
    if (level > 0) {
      if (s > 0 && t.keys[(size_t)id * F + s - 1] >= k) errs++;  // inner nodes in key order
    } else {
      // leaf entries are unordered but distinct
      for (uint32_t u = 0; u < count; ++u)
        if (u != (uint32_t)s && t.keys[(size_t)id * F + u] == k) errs++;
    }
    if (k < m.low || k >= m.high) errs++;
    if (level > 0) {
      if (s == 0 && k != m.low) errs++;
      if (v >= n_nodes) {
        errs++;
        continue;
      }
      const NodeMeta cm = t.meta[v];
      if (meta_level(cm.info) + 1 != level) errs++;
      if (cm.low != k) errs++;
      const uint64_t hi = ((uint32_t)s + 1 < count) ? t.keys[(size_t)id * F + s + 1] : m.high;
      if (cm.high != hi) errs++;
    }
  }
  if (lane_id() == 0) {
    if (m.next != NODE_NONE) {
      if (m.next >= n_nodes) errs++;
      else {
        const NodeMeta nm = t.meta[m.next];
        if (nm.low != m.high || meta_level(nm.info) != level) errs++;
      }
    } else if (m.high != KEY_EMPTY) {
      errs++;
    }
    if (level > 0 && count == 0) errs++;
  }
  for (int off = 32; off >= 1; off >>= 1) errs += __shfl_xor(errs, off, WAVE);
  if (lane_id() == 0 && errs) atomicAdd(bad, (unsigned long long)errs);
}

// ================================================================== launchers
template <int F>
static hipError_t launch_build_leaves_f(const TreeView& t, const uint64_t* ukey, const uint64_t* uval,
                                        uint32_t U, uint32_t fill, uint32_t L, uint64_t* lows,
                                        hipStream_t s) {
  const uint64_t n = (uint64_t)L * F;
  k_build_leaves<F><<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(t, ukey, uval, U, fill, L, lows);
  return hipGetLastError();
}
template <int F>
static hipError_t launch_build_inner_f(const TreeView& t, const uint64_t* lows_prev, uint32_t n_prev,
                                       uint32_t first_prev, uint32_t fill, uint32_t n_this,
                                       uint32_t first_this, uint32_t level, uint64_t* lows_out,
                                       hipStream_t s) {
  const uint64_t n = (uint64_t)n_this * F;
  k_build_inner<F><<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(t, lows_prev, n_prev, first_prev, fill,
                                                               n_this, first_this, level, lows_out);
  return hipGetLastError();
}

#define STM_DISPATCH_F(F_, CALL)   \
  switch (F_) {                    \
    case 32: { constexpr int F = 32; return CALL; } \
    case 64: { constexpr int F = 64; return CALL; } \
    case 128: { constexpr int F = 128; return CALL; } \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_build_leaves(const TreeView& t, const uint64_t* ukey, const uint64_t* uval,
                               uint32_t U, uint32_t fill, uint32_t L, uint64_t* lows, hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_build_leaves_f<F>(t, ukey, uval, U, fill, L, lows, s));
}
hipError_t launch_build_inner(const TreeView& t, const uint64_t* lows_prev, uint32_t n_prev,
                              uint32_t first_prev, uint32_t fill, uint32_t n_this,
                              uint32_t first_this, uint32_t level, uint64_t* lows_out,
                              hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_build_inner_f<F>(t, lows_prev, n_prev, first_prev, fill, n_this,
                                                   first_this, level, lows_out, s));
}

template <int F, int G>
static hipError_t launch_lookup_fg(const TreeView& t, const LookupArgs& a, hipStream_t s) {
  const uint64_t groups = (a.n_max + a.chunk - 1) / a.chunk;
  const uint64_t threads = groups * G;
  // Persistent grid: at most LOOKUP_GRID_CAP workgroups; each lane group walks one contiguous span.
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((threads + 255) / 256, LOOKUP_GRID_CAP);
  if (blocks == 0) return hipSuccess;
  if (a.route_hdr && a.index)
    k_lookup<F, G, true, true><<<blocks, 256, 0, s>>>(t, a.ukey, a.uop, a.gstart, a.npend, a.sidx,
                                                       a.counters, a.chunk, a.out_vals, a.out_found,
                                                       a.route_hdr, a.fences, a.lids, a.segs, a.desc,
                                                       a.ihint, a.gval, a.fanout,
                                                       a.fanout_cap);
  else if (a.route_hdr)
    k_lookup<F, G, true, false><<<blocks, 256, 0, s>>>(t, a.ukey, a.uop, a.gstart, a.npend, a.sidx,
                                                        a.counters, a.chunk, a.out_vals, a.out_found,
                                                        a.route_hdr, a.fences, a.lids, a.segs, a.desc,
                                                        a.ihint, a.gval, a.fanout,
                                                        a.fanout_cap);
  else
    k_lookup<F, G, false, false><<<blocks, 256, 0, s>>>(t, a.ukey, a.uop, a.gstart, a.npend, a.sidx,
                                                         a.counters, a.chunk, a.out_vals, a.out_found,
                                                         nullptr, nullptr, nullptr, nullptr, a.desc,
                                                         a.ihint, a.gval, a.fanout,
                                                         a.fanout_cap);
  return hipGetLastError();
}

template <int F>
static hipError_t launch_lookup_f(const TreeView& t, const LookupArgs& a, hipStream_t s) {
  switch (a.lanes) {
    case 8: return launch_lookup_fg<F, 8>(t, a, s);
    case 16: return launch_lookup_fg<F, 16>(t, a, s);
    case 32: return launch_lookup_fg<F, (F >= 32 ? 32 : 16)>(t, a, s);
    case 64: return launch_lookup_fg<F, (F >= 64 ? 64 : 32)>(t, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_lookup(const TreeView& t, const LookupArgs& a, hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_lookup_f<F>(t, a, s));
}

template <int F, int G>
static hipError_t launch_upsert_fast(const TreeView& t, const UpsertArgs& a, hipStream_t s,
                                     uint32_t& waves) {
  // Each wavefront takes an equal contiguous share of the merged INSERTs and works through it
  // in rounds of WAVE / G: a grid of exactly the wavefronts the chip holds at once finishes
  // in one generation of rounds instead of two.
  static std::atomic<int> resident{0};  // blocks of this instantiation the device holds at once
  if (resident.load(std::memory_order_relaxed) == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_upsert_fast<F, G>, 256, 0);
    if (e != hipSuccess) return e;
    resident.store(std::max(1, cus * per_cu), std::memory_order_relaxed);
  }
  const uint32_t blocks = std::min<uint32_t>(
      std::min<uint32_t>(a.blocks, (uint32_t)resident.load(std::memory_order_relaxed)),
      FAST_WAVES_MAX / 4);
  waves = blocks * 4;  // one scount entry per wavefront
  k_upsert_fast<F, G><<<blocks, 256, 0, s>>>(t, a.ukey, a.uval, a.uop, a.gstart, a.npend, a.sidx,
                                                a.ilist, a.ihint, a.counters, a.slist,
                                                a.scount, a.out_vals, a.out_found, a.desc,
                                                a.fanout, a.fanout_cap, a.gval);
  return hipGetLastError();
}

template <int F>
static hipError_t launch_upsert_f(const TreeView& t, const UpsertArgs& a, hipStream_t s) {
  hipError_t e = hipSuccess;
  const bool fast = a.ihint && a.ilist && a.slist && a.scount;
  uint32_t fast_waves = 0;
  if (fast) {
    switch (a.lanes) {
      case 8: e = launch_upsert_fast<F, (F >= 8 ? 8 : F)>(t, a, s, fast_waves); break;
      case 16: e = launch_upsert_fast<F, (F >= 16 ? 16 : F)>(t, a, s, fast_waves); break;
      case 32: e = launch_upsert_fast<F, (F >= 32 ? 32 : F)>(t, a, s, fast_waves); break;
      default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
  }
  k_upsert<F><<<a.blocks, 256, 0, s>>>(t, a.ukey, a.uval, a.uop, a.gstart, a.npend, a.sidx,
                                        fast ? a.slist : nullptr, a.scount, fast_waves,
                                        fast ? a.ihint : nullptr, a.counters, a.dlist,
                                        a.max_attempts, a.out_vals,
                                        a.out_found, a.desc);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_upsert_serial<F><<<1, 64, 0, s>>>(t, a.ukey, a.uval, a.gstart, a.npend, a.sidx, a.dlist,
                                      a.counters, a.out_vals, a.out_found, a.desc,
                                      a.host_counters);
  return hipGetLastError();
}

hipError_t launch_upsert(const TreeView& t, const UpsertArgs& a, hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_upsert_f<F>(t, a, s));
}

hipError_t launch_collect_leaves(const TreeView& t, uint32_t n_nodes, uint64_t* lows, uint32_t* ids,
                                 uint32_t* counter, hipStream_t s) {
  if (n_nodes == 0) return hipSuccess;
  k_collect_leaves<<<(n_nodes + 255) / 256, 256, 0, s>>>(t, n_nodes, lows, ids, counter);
  return hipGetLastError();
}
hipError_t launch_leaf_counts(const TreeView& t, const uint32_t* order, const uint32_t* ids,
                              uint32_t n_leaves, uint32_t* sorted_ids, uint32_t* counts,
                              hipStream_t s) {
  if (n_leaves == 0) return hipSuccess;
  k_leaf_counts<<<(n_leaves + 255) / 256, 256, 0, s>>>(t, order, ids, n_leaves, sorted_ids, counts);
  return hipGetLastError();
}
template <int F>
static hipError_t launch_dump_copy_f(const TreeView& t, const uint32_t* sorted_ids,
                                     const uint64_t* offsets, uint32_t n_leaves, uint64_t* out_k,
                                     uint64_t* out_v, hipStream_t s) {
  const uint64_t n = (uint64_t)n_leaves * F;
  if (n == 0) return hipSuccess;
  k_dump_copy<F><<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(t, sorted_ids, offsets, n_leaves, out_k, out_v);
  return hipGetLastError();
}
hipError_t launch_dump_copy(const TreeView& t, const uint32_t* sorted_ids, const uint64_t* offsets,
                            uint32_t n_leaves, uint64_t* out_k, uint64_t* out_v, hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_dump_copy_f<F>(t, sorted_ids, offsets, n_leaves, out_k, out_v, s));
}
template <int F>
static hipError_t launch_check_f(const TreeView& t, uint32_t n_nodes, unsigned long long* bad,
                                 hipStream_t s) {
  if (n_nodes == 0) return hipSuccess;
  const uint64_t threads = (uint64_t)n_nodes * WAVE;
  k_check<F><<<(uint32_t)((threads + 255) / 256), 256, 0, s>>>(t, n_nodes, bad);
  return hipGetLastError();
}
hipError_t launch_check(const TreeView& t, uint32_t n_nodes, unsigned long long* bad, hipStream_t s) {
  STM_DISPATCH_F(t.fanout, launch_check_f<F>(t, n_nodes, bad, s));
}

}  // namespace stm
