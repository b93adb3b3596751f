// stm_fold.hip — per-key fold of a sorted batch into one merged op per unique key.
//
// Reference pipeline (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:289-332):
//   cub::DeviceRunLengthEncode::Encode (:296) -> run count copied to the host (:303) ->
//   cub::DeviceScan::ExclusiveSum (:305) -> mergeMission (bpt-gpustm-kernel-aux.cuh:16-44)
//   -> tree kernel -> writebackAns (bpt-gpustm-kernel-aux.cuh:46-56).
// mergeMission walks each key group in input order: an op is answered by the value of
// the latest earlier INSERT of the group, otherwise by the tree; the merged op is
// INSERT(last inserted value) if the group has an insert, GET otherwise.
//
// Here one scan replaces RLE + scan + the thread-per-group loop, so skewed (Zipf) groups
// are spread over many threads and no run count travels to the host.  Because the groups
// are contiguous in the sorted order, the whole fold follows from three running values of a
// prefix: the number of group heads, the position of the last head and the position of the
// last INSERT.  For element i with head h(i), the latest earlier insert of its key is the
// prefix's last insert if it is >= h(i); the group's merged op is INSERT iff the last insert
// up to its end is >= its head.  A wavefront owns 512 consecutive elements as 8 rounds of
// 64 (coalesced loads and group outputs); within a round the three values of every lane's
// prefix come from two ballots (heads, inserts): a popcount and a highest set bit.  Three
// kernels: tile reduce, scan of the tile aggregates, apply.  Ops that the fold answers (an
// earlier insert of the same key exists) are written to the outputs here; the first op of
// each group up to and including its first insert waits for the tree ("npend" ops from
// "gstart").
#include "stm_common.hpp"
#include "stm_internal.hpp"

namespace stm {

constexpr int FOLD_BLOCK = 256;
constexpr int FOLD_WAVES = FOLD_BLOCK / WAVE;
constexpr int FOLD_ROUNDS = 8;                     // rounds of WAVE elements per wavefront
constexpr int FOLD_CHUNK = FOLD_ROUNDS * WAVE;     // elements of one wavefront
constexpr int FOLD_TILE = FOLD_WAVES * FOLD_CHUNK;  // elements of one block

struct Acc {
  uint32_t c;    // group heads in the range
  int32_t head;  // position of the last head (-1 if none)
  int32_t ins;   // position of the last INSERT (-1 if none)
};

STM_DEV Acc acc_identity() { return Acc{0u, -1, -1}; }
STM_DEV Acc acc_combine(const Acc& a, const Acc& b) {
  return Acc{a.c + b.c, max(a.head, b.head), max(a.ins, b.ins)};
}

// Position of the highest lane set in m within a round starting at rb, or `none`.
STM_DEV int32_t last_set(uint64_t m, uint32_t rb, int32_t none) {
  return m ? (int32_t)(rb + 63u - (uint32_t)__clzll((long long)m)) : none;
}

// With a batch descriptor the element count and the caller's arrays come from device memory
// (one launch sequence serves every batch size).
STM_DEV void fold_resolve(FoldArgs& a) {
  const BatchDesc* d = a.desc;
  if (!d) return;
  a.n = d->n;
  a.vals = d->vals;
  a.ops = d->ops;
  a.out_vals = d->out_vals;
  a.out_found = d->out_found;
}

// A wavefront's chunk: element r * WAVE + lane of the chunk in k[r], sx[r] of that lane.
struct Chunk {
  uint32_t c0;  // position of the chunk's first element
  uint64_t k[FOLD_ROUNDS];
  uint32_t sx[FOLD_ROUNDS];
  uint64_t pk, nk;  // keys just before and just after the chunk (when they exist)
};

STM_DEV void load_chunk(const FoldArgs& a, uint32_t c0, Chunk& ch) {
  const int lane = lane_id();
  ch.c0 = c0;
#pragma unroll
  for (int r = 0; r < FOLD_ROUNDS; ++r) {
    const uint32_t i = c0 + (uint32_t)r * WAVE + lane;
    ch.k[r] = i < a.n ? a.skeys[i] : 0;
    ch.sx[r] = i < a.n ? a.sidx[i] : 0;
  }
  ch.pk = (c0 > 0 && c0 <= a.n) ? a.skeys[c0 - 1] : 0;
  ch.nk = (c0 + FOLD_CHUNK < a.n) ? a.skeys[c0 + FOLD_CHUNK] : 0;
}

// Head and INSERT flags of round r for this lane, with the key after it.
STM_DEV void round_flags(const FoldArgs& a, const Chunk& ch, int r, bool& valid, bool& head,
                         bool& ins, uint64_t& next) {
  const int lane = lane_id();
  const uint32_t i = ch.c0 + (uint32_t)r * WAVE + lane;
  valid = i < a.n;
  const uint64_t up = shfl64(ch.k[r], lane > 0 ? lane - 1 : 0);
  const uint64_t prev_last = r > 0 ? readlane64(ch.k[r > 0 ? r - 1 : 0], WAVE - 1) : ch.pk;
  const uint64_t prev = lane > 0 ? up : prev_last;
  const uint64_t down = shfl64(ch.k[r], lane < WAVE - 1 ? lane + 1 : WAVE - 1);
  const uint64_t next_first = r + 1 < FOLD_ROUNDS ? readlane64(ch.k[r + 1 < FOLD_ROUNDS ? r + 1 : r], 0) : ch.nk;
  next = lane < WAVE - 1 ? down : next_first;
  head = valid && (i == 0 || ch.k[r] != prev);
  ins = valid && op_of(ch.sx[r]) == OP_INSERT;
}

// The chunk's aggregate (wave-uniform).
STM_DEV Acc chunk_acc(const FoldArgs& a, const Chunk& ch) {
  Acc t = acc_identity();
#pragma unroll
  for (int r = 0; r < FOLD_ROUNDS; ++r) {
    bool valid, head, ins;
    uint64_t next;
    round_flags(a, ch, r, valid, head, ins, next);
    const uint32_t rb = ch.c0 + (uint32_t)r * WAVE;
    const uint64_t hm = ballot(head), im = ballot(ins);
    t.c += (uint32_t)__popcll(hm);
    t.head = last_set(hm, rb, t.head);
    t.ins = last_set(im, rb, t.ins);
  }
  return t;
}

// Exclusive prefix over the block's wavefronts of their chunk aggregates, and the total.
STM_DEV Acc block_wave_prefix(const Acc& w, Acc* s_w /*[FOLD_WAVES]*/, Acc& total) {
  const int wave = threadIdx.x / WAVE;
  if (lane_id() == 0) s_w[wave] = w;
  __syncthreads();
  Acc before = acc_identity();
  total = acc_identity();
#pragma unroll
  for (int q = 0; q < FOLD_WAVES; ++q) {
    const Acc t = s_w[q];
    if (q < wave) before = acc_combine(before, t);
    total = acc_combine(total, t);
  }
  __syncthreads();
  return before;
}

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_reduce(FoldArgs a, Acc* tile_agg) {
  __shared__ Acc s_w[FOLD_WAVES];
  fold_resolve(a);
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  const uint32_t wave = threadIdx.x / WAVE;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    Chunk ch;
    load_chunk(a, b * FOLD_TILE + wave * FOLD_CHUNK, ch);
    Acc total;
    (void)block_wave_prefix(chunk_acc(a, ch), s_w, total);
    if (threadIdx.x == 0) tile_agg[b] = total;
  }
}

// Exclusive scan of one Acc per thread over the block; also returns the block total.
STM_DEV Acc block_exclusive_acc(const Acc& v, Acc* s_w /*[FOLD_WAVES]*/, Acc& total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  Acc x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const Acc y{(uint32_t)__shfl_up((int)x.c, off, WAVE), __shfl_up(x.head, off, WAVE),
                __shfl_up(x.ins, off, WAVE)};
    if (lane >= off) x = acc_combine(y, x);
  }
  Acc ex{(uint32_t)__shfl_up((int)x.c, 1, WAVE), __shfl_up(x.head, 1, WAVE), __shfl_up(x.ins, 1, WAVE)};
  if (lane == 0) ex = acc_identity();
  if (lane == WAVE - 1) s_w[wave] = x;
  __syncthreads();
  Acc wp = acc_identity();
  total = acc_identity();
#pragma unroll
  for (int w = 0; w < FOLD_WAVES; ++w) {
    const Acc t = s_w[w];
    if (w < wave) wp = acc_combine(wp, t);
    total = acc_combine(total, t);
  }
  __syncthreads();
  return acc_combine(wp, ex);
}

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_scan_tiles(Acc* tile_agg, FoldArgs a,
                                                                BatchCounters* counters) {
  __shared__ Acc s_w[FOLD_WAVES];
  fold_resolve(a);
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  const uint32_t per = (ntiles + FOLD_BLOCK - 1) / FOLD_BLOCK;
  const uint32_t b0 = min(threadIdx.x * per, ntiles);
  const uint32_t b1 = min(b0 + per, ntiles);
  Acc s = acc_identity();
  for (uint32_t b = b0; b < b1; ++b) s = acc_combine(s, tile_agg[b]);
  Acc total;
  Acc run = block_exclusive_acc(s, s_w, total);
  for (uint32_t b = b0; b < b1; ++b) {
    const Acc t = tile_agg[b];
    tile_agg[b] = run;  // becomes the exclusive prefix of tile b
    run = acc_combine(run, t);
  }
  if (threadIdx.x == 0) {
    // A sort that gave up a look-back wait may have mis-ordered the batch: the batch then has
    // no groups, so the fold, lookup, scan and upsert kernels do nothing and the tree keeps
    // its state from before the batch.
    const uint32_t stalled = a.sort_stall ? *a.sort_stall : 0u;
    counters->n_unique = stalled ? 0u : total.c;
    counters->n_upserts = 0;
    counters->n_deferred = 0;
    counters->queue_head = 0;
    counters->n_bad = 0;
    counters->n_fanout = 0;
    counters->sort_stall = stalled;
    counters->sort_fallback = a.sort_exact ? *a.sort_exact : 0u;
    counters->leaf_splits = 0;
  }
}

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_apply(FoldArgs a, const Acc* tile_prefix) {
  __shared__ Acc s_w[FOLD_WAVES];
  __shared__ uint32_t s_ins[FOLD_TILE];  // the tile's merged INSERT groups
  __shared__ uint32_t s_nins, s_base;
  fold_resolve(a);
  if (a.counters->sort_stall) return;  // the batch runs empty (k_fold_scan_tiles)
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  const uint32_t wave = threadIdx.x / WAVE;
  const uint64_t lt = lanemask_lt();
  const uint64_t le = lt | (1ull << lane_id());
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    if (threadIdx.x == 0) s_nins = 0;
    Chunk ch;
    load_chunk(a, b * FOLD_TILE + wave * FOLD_CHUNK, ch);
    Acc total;
    // State before this wavefront's chunk (block_wave_prefix syncs s_nins too).
    Acc carry = acc_combine(tile_prefix[b], block_wave_prefix(chunk_acc(a, ch), s_w, total));
    // The rounds in halves of HALF, each in three phases, so that the gathers of a half share
    // two round trips: the vector memory counter also counts stores, so a gather after a
    // round's output stores would wait for them, and one gather chain per round made the
    // kernel latency-bound.  (All eight rounds at once needed 138 registers.)
    constexpr int HALF = FOLD_ROUNDS / 2;
    constexpr uint32_t F_VALID = 1, F_INS = 2, F_END = 4;
#pragma unroll
    for (int h = 0; h < FOLD_ROUNDS; h += HALF) {
    // Phase 1, registers only: each round's flags, group index, head and latest earlier insert.
    uint32_t fl[HALF], gg[HALF];
    int32_t hd[HALF], lb[HALF];
#pragma unroll
    for (int q = 0; q < HALF; ++q) {
      const int r = h + q;
      bool valid, is_head, is_ins;
      uint64_t next;
      round_flags(a, ch, r, valid, is_head, is_ins, next);
      const uint32_t rb = ch.c0 + (uint32_t)r * WAVE;
      const uint32_t i = rb + lane_id();
      const uint64_t hm = ballot(is_head), im = ballot(is_ins);
      // This lane's prefix: heads up to and including it, its head, the last insert before it.
      const uint32_t c_incl = carry.c + (uint32_t)__popcll(hm & le);
      hd[q] = last_set(hm & le, rb, carry.head);
      const int32_t ins_before = last_set(im & lt, rb, carry.ins);
      lb[q] = ins_before >= hd[q] ? ins_before : -1;  // latest earlier insert of the key
      gg[q] = c_incl - 1;
      const bool group_end = valid && (i + 1 == a.n || next != ch.k[r]);
      fl[q] = (valid ? F_VALID : 0u) | (is_ins ? F_INS : 0u) | (group_end ? F_END : 0u);
      carry.c += (uint32_t)__popcll(hm);
      carry.head = last_set(hm, rb, carry.head);
      carry.ins = last_set(im, rb, carry.ins);
    }
    // Phase 2: input positions of the earlier inserts, then the values.  va = the latest
    // earlier insert's value (an op's answer, or the merged value of a group ending in a
    // non-insert), vb = the op's own value (the merged value of a group ending in an insert).
    uint32_t ia[HALF];
#pragma unroll
    for (int q = 0; q < HALF; ++q)
      ia[q] = ((fl[q] & F_VALID) && lb[q] >= 0) ? a.sidx[lb[q]] : 0u;
    uint64_t va[HALF], vb[HALF];
#pragma unroll
    for (int q = 0; q < HALF; ++q) {
      const int r = h + q;
      const bool valid = fl[q] & F_VALID, is_ins = fl[q] & F_INS, end = fl[q] & F_END;
      const bool real = ch.k[r] != KEY_EMPTY;
      const bool need_a = valid && lb[q] >= 0 && (a.out_vals || (end && !is_ins && real));
      const bool need_b = valid && end && is_ins && real;
      va[q] = need_a ? a.vals[pos_of(ia[q])] : 0ull;
      vb[q] = need_b ? a.vals[pos_of(ch.sx[r])] : 0ull;
    }
    // Phase 3: the outputs and the tile's insert list.
#pragma unroll
    for (int q = 0; q < HALF; ++q) {
      const int r = h + q;
      const bool valid = fl[q] & F_VALID, is_ins = fl[q] & F_INS, group_end = fl[q] & F_END;
      const int32_t i = (int32_t)(ch.c0 + (uint32_t)r * WAVE + lane_id());
      const int32_t head = hd[q];
      const int32_t last_before = lb[q];
      const uint32_t g = gg[q];
      bool append = false;
      if (valid) {
        const uint64_t key = ch.k[r];
        if (a.lastins) a.lastins[i] = is_ins ? i : last_before;
        if (a.out_vals && last_before >= 0) {
          const uint32_t pos = pos_of(ch.sx[r]);
          a.out_vals[pos] = va[q];
          a.out_found[pos] = 1;
        }
        if (a.gstart && is_ins && last_before < 0) a.npend[g] = (uint32_t)(i - head + 1);
        if (group_end) {
          a.ukey[g] = key;
          if (a.gstart) a.gstart[g] = (uint32_t)head;
          // the group's last insert: this op if it is one, else the latest earlier one
          const int32_t last_in = is_ins ? i : last_before;
          uint8_t op;
          if (key == KEY_EMPTY) {
            op = UOP_SKIP;
            atomicAdd(&a.counters->n_bad, (uint32_t)(i - head + 1));
          } else if (last_in >= 0) {
            op = OP_INSERT;
            a.uval[g] = is_ins ? vb[q] : va[q];
            append = true;
          } else {
            op = OP_GET;
            if (a.gstart) a.npend[g] = (uint32_t)(i - head + 1);
          }
          a.uop[g] = op;
        }
      }
      const uint64_t want = ballot(append);
      uint32_t wb = 0;
      if (lane_id() == 0 && want) wb = atomicAdd(&s_nins, (uint32_t)__popcll(want));  // LDS
      wb = (uint32_t)__shfl((int)wb, 0, WAVE);
      if (append) s_ins[wb + (uint32_t)__popcll(want & lt)] = g;
    }
    }  // halves
    // One counter update per tile (a hot atomic per wavefront round would serialize the
    // kernel on one L2 channel), then the tile's INSERT groups go to the batch's list.
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_nins ? atomicAdd(&a.counters->n_upserts, s_nins) : 0u;
    __syncthreads();
    if (a.ilist)
      for (uint32_t k = threadIdx.x; k < s_nins; k += FOLD_BLOCK) a.ilist[s_base + k] = s_ins[k];
    __syncthreads();
  }  // tiles
}

size_t fold_scratch_bytes(uint64_t n) {
  uint64_t ntiles = (n + FOLD_TILE - 1) / FOLD_TILE;
  if (ntiles == 0) ntiles = 1;
  return (size_t)ntiles * sizeof(Acc);
}

// a.n is the element count, or with a.desc the largest count the scratch was sized for.
hipError_t fold_batch(const FoldArgs& a, hipStream_t stream) {
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  if (a.n == 0) return hipSuccess;
  const uint32_t grid = std::min<uint32_t>(ntiles, 2048);
  Acc* tiles = (Acc*)a.tile_scratch;
  k_fold_reduce<<<grid, FOLD_BLOCK, 0, stream>>>(a, tiles);
  k_fold_scan_tiles<<<1, FOLD_BLOCK, 0, stream>>>(tiles, a, a.counters);
  k_fold_apply<<<grid, FOLD_BLOCK, 0, stream>>>(a, tiles);
  return hipGetLastError();
}

}  // namespace stm
