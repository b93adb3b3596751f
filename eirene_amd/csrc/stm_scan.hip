// stm_scan.hip — range scans with exact sequential semantics inside a batch.
//
// No reference counterpart: the reference declares RANGE (op 3, src/input-helper.h:14-17)
// but never implements it (src/tree.h:37-39, src/main.cpp:30-34), and its v38 leaves keep no
// sibling links (src/gpu-stm-tree-v38/bpt-gpustm-kernel-tree.cuh:413-445).  The v35 variant
// maintains right-sibling pointers on split (src/gpu-stm-tree-v35/bpt-gpustm-kernel-tx.cuh:436-444);
// this tree keeps them in NodeMeta::next.
//
// SCAN(k, L) at input position i returns the first L pairs (x, v) with x >= k in the state
// before op i.  That state is the tree before the batch plus the batch's inserts at input
// positions < i.  Scans therefore run after the per-key fold and before the upsert kernel:
//   * the tree part is read from the unmodified pre-batch tree (leaf, then right siblings);
//   * the batch part is read from the fold's output, one group per unique key in key order:
//     a group contributes key x iff its first insert precedes i; its value is the latest
//     insert of x before i, found by a binary search over the group's input positions
//     (stable sort keeps them increasing) and the fold's per-entry "latest insert" index.
// One wavefront merges the two sorted streams 64 candidates at a time: each round takes the
// next leaf chunk and the next 64 groups, finalises every candidate not above the smaller of
// the two chunks' last keys (batch values shadow equal tree keys), places each by its merge
// rank (lane binary searches over the other chunk) and writes up to L - written pairs.
//
// Output layout: scan op i's pairs start at off_i = the sum of L over the scan ops before it
// in input order (three kernels compute off_i); out_vals[i] = pairs written, out_found[i] =
// (pairs > 0), written after the upsert kernel (which answers every non-insert op of a key
// group, scans included, with the key's value).
#include <algorithm>

#include "stm_common.hpp"
#include "stm_internal.hpp"
#include "stm_tree.hpp"

namespace stm {

constexpr int OFF_BLOCK = 256;
constexpr int OFF_ITEMS = 16;
constexpr int OFF_TILE = OFF_BLOCK * OFF_ITEMS;
constexpr uint32_t SCAN_GRID_CAP = 2048;
// Wavefronts that share the scans of one 64-op window.  A scan is a chain of dependent leaf
// reads, so a window holding several scans set the kernel's time when one wavefront ran them
// all.
constexpr uint32_t SCAN_SPLIT = 4;
constexpr uint32_t SCAN_ROUND_CAP = 1u << 24;

STM_DEV uint64_t scan_len(uint8_t op, uint64_t v) {
  return (op == OP_RANGE && v >= 1 && v <= SCAN_MAX_LEN) ? v : 0;
}

STM_DEV uint64_t block_exclusive_sum_u64(uint64_t v, uint64_t* s_w /*[4]*/, uint64_t& total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const uint64_t y = shfl64(x, lane - off < 0 ? lane : lane - off);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) s_w[wave] = x;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < OFF_BLOCK / WAVE; ++w) {
    const uint64_t t = s_w[w];
    if (w < wave) before += t;
    total += t;
  }
  __syncthreads();
  return before + x - v;
}

struct ScanIn {
  const uint64_t* keys;
  const uint64_t* vals;
  const uint8_t* ops;
  uint32_t n;
  uint64_t* out_vals;
  uint8_t* out_found;
  uint64_t* scan_keys;
  uint64_t* scan_vals;
  uint64_t scan_cap;
};

STM_DEV ScanIn scan_in(const ScanArgs& a) {
  const BatchDesc* d = a.desc;
  if (d) return ScanIn{d->keys, d->vals, d->ops, d->n, d->out_vals, d->out_found, d->scan_keys,
                       d->scan_vals, d->scan_cap};
  return ScanIn{a.keys, a.vals, a.ops, a.n, a.out_vals, a.out_found, a.scan_keys, a.scan_vals,
                a.scan_cap};
}

// Tile sums of the scan lengths in input order.
__global__ __launch_bounds__(OFF_BLOCK) void k_scanoff_reduce(ScanArgs a) {
  __shared__ uint64_t s_w[OFF_BLOCK / WAVE];
  const ScanIn in = scan_in(a);
  const uint32_t ntiles = (in.n + OFF_TILE - 1) / OFF_TILE;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    uint64_t acc = 0;
    for (int r = 0; r < OFF_ITEMS; ++r) {
      const uint64_t i = (uint64_t)b * OFF_TILE + (uint64_t)r * OFF_BLOCK + threadIdx.x;
      if (i < in.n) {
        const uint8_t op = in.ops[i];
        if (op == OP_RANGE) acc += scan_len(op, in.vals[i]);
      }
    }
    uint64_t total;
    (void)block_exclusive_sum_u64(acc, s_w, total);
    if (threadIdx.x == 0) a.tile_sums[b] = total;
  }
}

// Exclusive scan of the tile sums (one block); the batch total goes to the counters.
__global__ __launch_bounds__(OFF_BLOCK) void k_scanoff_tiles(ScanArgs a) {
  __shared__ uint64_t s_w[OFF_BLOCK / WAVE];
  const ScanIn in = scan_in(a);
  const uint32_t ntiles = (in.n + OFF_TILE - 1) / OFF_TILE;
  const uint32_t per = (ntiles + OFF_BLOCK - 1) / OFF_BLOCK;
  const uint32_t b0 = min(threadIdx.x * per, ntiles);
  const uint32_t b1 = min(b0 + per, ntiles);
  uint64_t s = 0;
  for (uint32_t b = b0; b < b1; ++b) s += a.tile_sums[b];
  uint64_t total;
  uint64_t run = block_exclusive_sum_u64(s, s_w, total);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint64_t t = a.tile_sums[b];
    a.tile_sums[b] = run;
    run += t;
  }
  if (threadIdx.x == 0) {
    a.counters->scan_total = total;
    a.counters->scan_bad = 0;
  }
}

// off_i for every scan op (input order), into soff[i].
__global__ __launch_bounds__(OFF_BLOCK) void k_scanoff_apply(ScanArgs a) {
  __shared__ uint64_t s_w[OFF_BLOCK / WAVE];
  const ScanIn in = scan_in(a);
  const uint32_t ntiles = (in.n + OFF_TILE - 1) / OFF_TILE;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    uint64_t carry = a.tile_sums[b];
    for (int r = 0; r < OFF_ITEMS; ++r) {
      const uint64_t r0 = (uint64_t)b * OFF_TILE + (uint64_t)r * OFF_BLOCK;
      if (r0 >= in.n) break;  // uniform over the block
      const uint64_t i = r0 + threadIdx.x;
      uint8_t op = 0xff;
      uint64_t len = 0;
      if (i < in.n) {
        op = in.ops[i];
        if (op == OP_RANGE) len = scan_len(op, in.vals[i]);
      }
      uint64_t total;
      const uint64_t ex = block_exclusive_sum_u64(len, s_w, total);
      if (op == OP_RANGE) a.soff[i] = carry + ex;
      carry += total;
    }
  }
}

// Number of lanes j < n whose (ascending) value v_j is below x; every lane runs the same
// steps, so the shuffles never read from an inactive lane.
STM_DEV uint32_t lanes_below(uint64_t v, uint32_t n, uint64_t x) {
  uint32_t pos = 0;
#pragma unroll
  for (int step = 64; step >= 1; step >>= 1) {
    const uint32_t probe = min(pos + (uint32_t)step - 1, 63u);
    const uint64_t pv = shfl64(v, (int)probe);
    if (pos + (uint32_t)step <= n && pv < x) pos += step;
  }
  return pos;
}

STM_DEV uint64_t mask_below(uint32_t r) { return r >= 64 ? ~0ull : ((1ull << r) - 1); }

template <int F>
STM_DEV uint32_t wave_find_leaf(const TreeView& t, uint64_t key) {
  const uint32_t cap = t.capacity;
  uint32_t node = (uint32_t)min<uint64_t>(uniform64(t.hdr->root), cap - 1);
  const uint32_t height = (uint32_t)min<uint64_t>(uniform64(t.hdr->height), 64);
  const int lane = lane_id();
  for (uint32_t lvl = height; lvl > 1; --lvl) {
    uint32_t cnt = 0;
#pragma unroll
    for (int s0 = 0; s0 < F; s0 += WAVE) {
      const int s = s0 + lane;
      const bool le = s < F && t.keys[(uint64_t)node * F + s] <= key;
      cnt += (uint32_t)__popcll(ballot(le));
    }
    const uint32_t slot = cnt ? cnt - 1 : 0;
    node = (uint32_t)min<uint64_t>(uniform64(t.vals[(uint64_t)node * F + slot]), cap - 1);
  }
  return node;
}

template <int F>
STM_DEV uint32_t wave_leaf_below(const TreeView& t, uint32_t leaf, uint64_t key) {
  const int lane = lane_id();
  uint32_t cnt = 0;
#pragma unroll
  for (int s0 = 0; s0 < F; s0 += WAVE) {
    const int s = s0 + lane;
    const bool lt = s < F && t.keys[(uint64_t)leaf * F + s] < key;
    cnt += (uint32_t)__popcll(ballot(lt));
  }
  return cnt;
}

// The entries of leaf `leaf` (the first `count` slots; a leaf's entries are unordered) in key
// order: position p = j * WAVE + lane holds key sk[j] from slot ss[j] (KEY_EMPTY past count).
// A bitonic sorting network over the wavefront's N = WAVE * SPL elements (element j * WAVE +
// lane in register j): 21 compare-exchange steps for N = 64, 28 for 128, each a lane shuffle of
// the key and slot (or, at distance 64, the lane's other register).  The keys are distinct and
// KEY_EMPTY pads past count, so the order is exact.  (Ranking each entry with two wave-uniform
// passes over the entries took a serving batch's scan kernel from 49 to 115 µs.)
template <int F>
STM_DEV void leaf_in_key_order(const TreeView& t, uint32_t leaf, uint32_t count,
                               uint64_t (&sk)[(F + WAVE - 1) / WAVE],
                               uint32_t (&ss)[(F + WAVE - 1) / WAVE]) {
  constexpr int SPL = (F + WAVE - 1) / WAVE;
  constexpr uint32_t N = (uint32_t)(WAVE * SPL);
  static_assert(SPL == 1 || SPL == 2, "one or two slots per lane");
  const uint32_t lane = (uint32_t)lane_id();
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const uint32_t sl = (uint32_t)j * WAVE + lane;
    sk[j] = sl < count ? t.keys[(uint64_t)leaf * F + sl] : KEY_EMPTY;
    ss[j] = sl;
  }
#pragma unroll
  for (uint32_t size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (uint32_t d = size >> 1; d > 0; d >>= 1) {
      if (d >= (uint32_t)WAVE) {
        // SPL == 2, d == 64: elements lane and 64 + lane, both in this lane
        if constexpr (SPL == 2) {
          const bool up = (lane & size) == 0;  // ascending block (size >= 128 here)
          const bool sw = up ? sk[1] < sk[0] : sk[0] < sk[1];
          const uint64_t k0 = sk[0];
          const uint32_t s0 = ss[0];
          if (sw) {
            sk[0] = sk[1];
            ss[0] = ss[1];
            sk[1] = k0;
            ss[1] = s0;
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < SPL; ++j) {
          const uint32_t e = (uint32_t)j * WAVE + lane;
          const uint64_t pk = shfl64(sk[j], (int)(lane ^ d));
          const uint32_t ps = (uint32_t)__shfl((int)ss[j], (int)(lane ^ d), WAVE);
          const bool up = (e & size) == 0;
          const bool lower = (e & d) == 0;
          // the lower element of an ascending pair keeps the smaller key, and so on
          const bool take = (lower == up) ? pk < sk[j] : pk > sk[j];
          if (take) {
            sk[j] = pk;
            ss[j] = ps;
          }
        }
      }
    }
  }
}

// Entry at key-order position p (lane-dependent, p < F) of a leaf_in_key_order view.
template <int F, typename T>
STM_DEV T at_position(const T (&x)[(F + WAVE - 1) / WAVE], uint32_t p) {
  T r = x[0];
#pragma unroll
  for (int j = 0; j < (F + WAVE - 1) / WAVE; ++j) {
    T v;
    if constexpr (sizeof(T) == 8) v = shfl64(x[j], (int)(p & 63));
    else v = (T)__shfl((int)x[j], (int)(p & 63), WAVE);
    if ((p >> 6) == (uint32_t)j) r = v;
  }
  return r;
}

// One scan op, run by the whole wavefront; returns the pairs written.
template <int F>
STM_DEV uint32_t run_scan(const TreeView& t, const ScanArgs& a, const ScanIn& in, uint32_t U,
                          uint64_t k, uint32_t i, uint32_t L, uint64_t off) {
  const int lane = lane_id();
  const uint32_t cap = t.capacity;
  constexpr int SPL = (F + WAVE - 1) / WAVE;
  uint32_t tl = wave_find_leaf<F>(t, k);
  uint32_t tcount = min(meta_count(uniform64(t.meta[tl].info)), (uint32_t)F);
  uint64_t tsk[SPL];  // the current leaf in key order
  uint32_t tss[SPL];
  leaf_in_key_order<F>(t, tl, tcount, tsk, tss);
  uint32_t ts = wave_leaf_below<F>(t, tl, k);
  bool t_done = false;
  // First group with key >= k (groups are in ascending key order).
  uint32_t bg;
  {
    uint32_t lo = 0, hi = U;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.ukey[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    bg = uniform32(lo);
  }
  uint32_t cnt = 0;
  // Every round consumes a whole chunk of one stream; the caps only guard a corrupt pool.
  for (uint32_t round = 0; cnt < L && round < SCAN_ROUND_CAP; ++round) {
    // Next non-empty leaf chunk.
    for (uint32_t hop = 0; !t_done && ts >= tcount; ++hop) {
      const uint64_t nx = uniform64(t.meta[tl].next);
      if (nx == NODE_NONE || nx >= cap || hop >= SCAN_ROUND_CAP) {
        t_done = true;
      } else {
        tl = (uint32_t)nx;
        tcount = min(meta_count(uniform64(t.meta[tl].info)), (uint32_t)F);
        leaf_in_key_order<F>(t, tl, tcount, tsk, tss);
        ts = 0;
      }
    }
    if (t_done && bg >= U) break;
    const uint32_t tn = t_done ? 0u : min(tcount - ts, (uint32_t)WAVE);
    const bool tv = (uint32_t)lane < tn;
    // Key-order position ts + lane, read from every lane (the shuffles need their sources
    // active) and then selected.
    const uint32_t tp = min(ts + (uint32_t)lane, (uint32_t)F - 1);
    const uint64_t tk_at = at_position<F>(tsk, tp);
    const uint64_t tslot = (uint64_t)tl * F + at_position<F>(tss, tp);
    const uint64_t tk = tv ? tk_at : KEY_EMPTY;
    const bool t_more = !t_done && (ts + tn < tcount || uniform64(t.meta[tl].next) != NODE_NONE);
    const uint64_t lastT = t_more ? readlane64(tk, (int)tn - 1) : KEY_EMPTY;

    const uint32_t bn = bg < U ? min(U - bg, (uint32_t)WAVE) : 0u;
    const bool bv = (uint32_t)lane < bn;
    uint64_t bk = KEY_EMPTY;
    uint64_t bval = 0;
    bool vis = false;
    if (bv) {
      const uint32_t g = bg + lane;
      bk = a.ukey[g];
      if (a.uop[g] == OP_INSERT) {
        const uint32_t gs = a.gstart[g];
        const uint32_t fi = gs + max(a.npend[g], 1u) - 1;  // sorted index of the first insert
        const uint32_t ge = (g + 1 < U) ? a.gstart[g + 1] : in.n;
        if (pos_of(a.sidx[fi]) < i) {
          uint32_t lo = fi, hi = ge;  // sidx[lo] < i; the last such entry lies in [lo, hi)
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pos_of(a.sidx[mid]) < i) lo = mid;
            else hi = mid;
          }
          const int32_t li = a.lastins[lo];
          if (li >= 0) {
            vis = true;
            bval = in.vals[pos_of(a.sidx[li])];
          }
        }
      }
    }
    const bool b_more = bg + bn < U;
    const uint64_t lastB = b_more ? readlane64(bk, (int)bn - 1) : KEY_EMPTY;
    const uint64_t bound = min(lastT, lastB);

    const bool t_el = tv && tk <= bound;
    const bool b_raw = bv && bk <= bound;
    const bool b_el = b_raw && vis;
    const uint64_t visM = ballot(b_el);
    const uint32_t rB = lanes_below(bk, bn, tk);  // groups below this tree key
    const uint64_t bk_at = shfl64(bk, (int)min(rB, 63u));
    const bool shadow = t_el && rB < bn && bk_at == tk && ((visM >> rB) & 1ull);
    const uint64_t shM = ballot(shadow);
    const uint64_t tM = ballot(t_el);
    const uint32_t rT = lanes_below(tk, tn, bk);  // tree keys below this group key
    const uint64_t lt = lanemask_lt();
    const uint32_t room = L - cnt;
    if (t_el && !shadow) {
      const uint32_t p = (uint32_t)lane - (uint32_t)__popcll(shM & lt) +
                         (uint32_t)__popcll(visM & mask_below(rB));
      if (p < room) {
        in.scan_keys[off + cnt + p] = tk;
        in.scan_vals[off + cnt + p] = t.vals[tslot];
      }
    }
    if (b_el) {
      const uint32_t p = (uint32_t)__popcll(visM & lt) + rT - (uint32_t)__popcll(shM & mask_below(rT));
      if (p < room) {
        in.scan_keys[off + cnt + p] = bk;
        in.scan_vals[off + cnt + p] = bval;
      }
    }
    const uint32_t emitted = (uint32_t)(__popcll(tM) - __popcll(shM) + __popcll(visM));
    cnt += min(emitted, room);
    ts += (uint32_t)__popcll(tM);
    bg += (uint32_t)__popcll(ballot(b_raw));
  }
  return cnt;
}

// One wavefront per window of 64 input positions: the window's scan ops run one after the
// other, each by the whole wavefront.  soff[i] becomes the pair count of scan op i.
template <int F>
__global__ __launch_bounds__(256) void k_range_scan(TreeView t, ScanArgs a) {
  const ScanIn in = scan_in(a);
  const uint32_t U = a.counters->n_unique;
  const int lane = lane_id();
  const uint64_t wave0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t nwin = ((uint64_t)in.n + WAVE - 1) / WAVE;
  const uint64_t lt = lanemask_lt();
  // Work item p = (window p / SCAN_SPLIT, share p % SCAN_SPLIT): the wavefront runs the scans
  // of the window whose rank among the window's scans is p % SCAN_SPLIT modulo SCAN_SPLIT, so
  // a window with several scans spreads them over SCAN_SPLIT wavefronts.
  for (uint64_t p = wave0; p < nwin * SCAN_SPLIT; p += nwaves) {
    const uint64_t w = p / SCAN_SPLIT;
    const uint32_t share = (uint32_t)(p % SCAN_SPLIT);
    const uint64_t i_l = w * WAVE + lane;
    const bool is_scan = i_l < in.n && in.ops[i_l] == OP_RANGE;
    const uint32_t rank = (uint32_t)__popcll(ballot(is_scan) & lt);
    const bool mine = is_scan && rank % SCAN_SPLIT == share;
    uint64_t todo = ballot(mine);
    uint64_t k_l = 0, v_l = 0, off_l = 0;
    if (mine) {
      k_l = in.keys[i_l];
      v_l = in.vals[i_l];
      off_l = a.soff[i_l];
    }
    while (todo) {
      const int src = __builtin_ctzll(todo);
      todo &= todo - 1;
      const uint64_t k = readlane64(k_l, src);
      const uint64_t L = readlane64(v_l, src);
      const uint64_t off = readlane64(off_l, src);
      const uint32_t i = (uint32_t)(w * WAVE + src);
      uint32_t c = 0;
      const bool ok = L >= 1 && L <= SCAN_MAX_LEN && off + L <= in.scan_cap && k != KEY_EMPTY;
      if (ok) {
        c = run_scan<F>(t, a, in, U, k, i, (uint32_t)L, off);
      } else if (lane == 0) {
        atomicAdd(&a.counters->scan_bad, 1u);
      }
      if (lane == src) a.soff[i_l] = c;
    }
  }
}

// After the upsert kernel: the scan ops' answers.
__global__ __launch_bounds__(256) void k_scan_answers(ScanArgs a) {
  const ScanIn in = scan_in(a);
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < in.n; i += (uint64_t)gridDim.x * 256ull) {
    if (in.ops[i] == OP_RANGE) {
      const uint64_t c = a.soff[i];
      in.out_vals[i] = c;
      in.out_found[i] = c ? 1 : 0;
    }
  }
}

size_t scan_tile_words(uint64_t n) {
  uint64_t t = (n + OFF_TILE - 1) / OFF_TILE;
  return (size_t)std::max<uint64_t>(t, 1);
}

hipError_t launch_scan_offsets(const ScanArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const uint32_t ntiles = (a.n + OFF_TILE - 1) / OFF_TILE;
  const uint32_t grid = std::min<uint32_t>(ntiles, SCAN_GRID_CAP);
  k_scanoff_reduce<<<grid, OFF_BLOCK, 0, s>>>(a);
  k_scanoff_tiles<<<1, OFF_BLOCK, 0, s>>>(a);
  k_scanoff_apply<<<grid, OFF_BLOCK, 0, s>>>(a);
  return hipGetLastError();
}

template <int F>
static hipError_t launch_range_scan_f(const TreeView& t, const ScanArgs& a, hipStream_t s) {
  const uint64_t waves = ((uint64_t)a.n + WAVE - 1) / WAVE * SCAN_SPLIT;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((waves + 3) / 4, SCAN_GRID_CAP));
  k_range_scan<F><<<blocks, 256, 0, s>>>(t, a);
  return hipGetLastError();
}

hipError_t launch_range_scan(const TreeView& t, const ScanArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  switch (t.fanout) {
    case 32: return launch_range_scan_f<32>(t, a, s);
    case 64: return launch_range_scan_f<64>(t, a, s);
    case 128: return launch_range_scan_f<128>(t, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_scan_answers(const ScanArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)a.n + 255) / 256, 4096));
  k_scan_answers<<<blocks, 256, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace stm
