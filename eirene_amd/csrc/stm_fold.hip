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
// Here one segmented scan replaces RLE + scan + the thread-per-group loop, so skewed
// (Zipf) groups are spread over many threads and no run count travels to the host.
// The scan state of a prefix is (has_head, last head position, first and last insert
// position since that head, number of heads).  Three kernels: tile reduce, scan of the
// tile aggregates, apply.  Ops that the fold answers (an earlier insert of the same key
// exists) are written to the outputs here; the first op of each group up to and
// including its first insert waits for the tree ("npend" ops from "gstart").
#include "stm_common.hpp"
#include "stm_internal.hpp"

namespace stm {

constexpr int FOLD_BLOCK = 256;
constexpr int FOLD_ITEMS = 16;
constexpr int FOLD_TILE = FOLD_BLOCK * FOLD_ITEMS;
constexpr int FOLD_WAVES = FOLD_BLOCK / WAVE;

struct Seg {
  int32_t h;      // 1 if the range contains a group head
  int32_t head;   // position of the last head in the range (-1 if none)
  int32_t first;  // first insert after the last head (or in the whole range if h == 0)
  int32_t last;   // last insert after the last head (or in the whole range if h == 0)
  uint32_t c;     // heads in the range
};

STM_DEV Seg seg_identity() { return Seg{0, -1, -1, -1, 0u}; }

STM_DEV Seg seg_combine(const Seg& a, const Seg& b) {
  Seg r;
  r.h = a.h | b.h;
  r.head = b.h ? b.head : a.head;
  r.first = b.h ? b.first : (a.first >= 0 ? a.first : b.first);
  r.last = b.h ? b.last : max(a.last, b.last);
  r.c = a.c + b.c;
  return r;
}

STM_DEV Seg seg_shfl_up(const Seg& s, int off) {
  Seg r;
  r.h = __shfl_up(s.h, off, WAVE);
  r.head = __shfl_up(s.head, off, WAVE);
  r.first = __shfl_up(s.first, off, WAVE);
  r.last = __shfl_up(s.last, off, WAVE);
  r.c = __shfl_up(s.c, off, WAVE);
  return r;
}

// Exclusive scan of one Seg per thread over the block; also returns the block total.
STM_DEV Seg block_exclusive_seg(const Seg& v, Seg* s_w /*[FOLD_WAVES]*/, Seg& total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  Seg x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    Seg y = seg_shfl_up(x, off);
    if (lane >= off) x = seg_combine(y, x);
  }
  Seg ex = seg_shfl_up(x, 1);
  if (lane == 0) ex = seg_identity();
  if (lane == WAVE - 1) s_w[wave] = x;
  __syncthreads();
  Seg wp = seg_identity();
  total = seg_identity();
#pragma unroll
  for (int w = 0; w < FOLD_WAVES; ++w) {
    Seg t = s_w[w];
    if (w < wave) wp = seg_combine(wp, t);
    total = seg_combine(total, t);
  }
  __syncthreads();
  return seg_combine(wp, ex);
}

STM_DEV Seg seg_element(const FoldArgs& a, uint32_t i, uint64_t key, uint64_t prev_key) {
  const bool head = (i == 0) || key != prev_key;
  const bool ins = op_of(a.sidx[i]) == OP_INSERT;  // op code packed by the sort
  Seg e;
  e.h = head ? 1 : 0;
  e.head = head ? (int32_t)i : -1;
  e.first = ins ? (int32_t)i : -1;
  e.last = e.first;
  e.c = head ? 1u : 0u;
  return e;
}

// A tile is processed in FOLD_ITEMS rounds of FOLD_BLOCK consecutive elements: thread t
// owns element round * FOLD_BLOCK + t, so every global access of a round is coalesced and
// the group outputs of a round land in consecutive slots.
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

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_reduce(FoldArgs a, Seg* tile_agg) {
  __shared__ Seg s_w[FOLD_WAVES];
  fold_resolve(a);
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const uint32_t tile0 = b * FOLD_TILE;
    Seg acc = seg_identity();
    for (int r = 0; r < FOLD_ITEMS; ++r) {
      const uint32_t r0 = tile0 + r * FOLD_BLOCK;
      if (r0 >= a.n) break;  // uniform over the block
      const uint32_t i = r0 + threadIdx.x;
      Seg e = seg_identity();
      if (i < a.n) e = seg_element(a, i, a.skeys[i], i ? a.skeys[i - 1] : 0);
      Seg total;
      (void)block_exclusive_seg(e, s_w, total);
      acc = seg_combine(acc, total);
    }
    if (threadIdx.x == 0) tile_agg[b] = acc;
  }
}

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_scan_tiles(Seg* tile_agg, FoldArgs a,
                                                                BatchCounters* counters) {
  __shared__ Seg s_w[FOLD_WAVES];
  fold_resolve(a);
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  const uint32_t per = (ntiles + FOLD_BLOCK - 1) / FOLD_BLOCK;
  const uint32_t b0 = threadIdx.x * per;
  const uint32_t b1 = min(b0 + per, ntiles);
  Seg s = seg_identity();
  for (uint32_t b = b0; b < b1; ++b) s = seg_combine(s, tile_agg[b]);
  Seg total;
  Seg run = block_exclusive_seg(s, s_w, total);
  for (uint32_t b = b0; b < b1; ++b) {
    Seg t = tile_agg[b];
    tile_agg[b] = run;  // becomes the exclusive prefix of tile b
    run = seg_combine(run, t);
  }
  if (threadIdx.x == 0) {
    counters->n_unique = total.c;
    counters->n_upserts = 0;
    counters->n_deferred = 0;
    counters->queue_head = 0;
    counters->n_bad = 0;
    counters->n_slow = 0;
  }
}

__global__ __launch_bounds__(FOLD_BLOCK) void k_fold_apply(FoldArgs a, const Seg* tile_prefix) {
  __shared__ Seg s_w[FOLD_WAVES];
  __shared__ uint32_t s_ins[FOLD_TILE];  // the tile's merged INSERT groups
  __shared__ uint32_t s_nins, s_base;
  fold_resolve(a);
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
  const uint32_t tile0 = b * FOLD_TILE;
  if (threadIdx.x == 0) s_nins = 0;
  __syncthreads();
  Seg carry = tile_prefix[b];  // state of every element before this round
  for (int r = 0; r < FOLD_ITEMS; ++r) {
    const uint32_t r0 = tile0 + r * FOLD_BLOCK;
    if (r0 >= a.n) break;  // uniform over the block
    const uint32_t i = r0 + threadIdx.x;
    const bool valid = i < a.n;
    uint64_t k = 0, next = 0;
    Seg e = seg_identity();
    if (valid) {
      k = a.skeys[i];
      next = (i + 1 < a.n) ? a.skeys[i + 1] : k;
      e = seg_element(a, i, k, i ? a.skeys[i - 1] : 0);
    }
    Seg total;
    const Seg before = seg_combine(carry, block_exclusive_seg(e, s_w, total));
    carry = seg_combine(carry, total);
    bool append = false;
    uint32_t g = 0;
    if (valid) {
      const int32_t last_before = e.h ? -1 : before.last;  // latest earlier insert of this key
      if (a.lastins) a.lastins[i] = e.first >= 0 ? (int32_t)i : last_before;
      if (a.out_vals && last_before >= 0) {
        const uint32_t pos = pos_of(a.sidx[i]);
        a.out_vals[pos] = a.vals[pos_of(a.sidx[last_before])];
        a.out_found[pos] = 1;
      }
      const Seg x = seg_combine(before, e);
      const bool group_end = (i + 1 == a.n) || next != k;
      if (group_end) {
        g = x.c - 1;
        a.ukey[g] = k;
        if (a.gstart) {
          a.gstart[g] = (uint32_t)x.head;
          a.npend[g] = (uint32_t)((x.first >= 0 ? x.first : (int32_t)i) - x.head + 1);
        }
        uint8_t op;
        if (k == KEY_EMPTY) {
          op = UOP_SKIP;
          atomicAdd(&a.counters->n_bad, (uint32_t)(i - x.head + 1));
        } else if (x.last >= 0) {
          op = OP_INSERT;
          a.uval[g] = a.vals[pos_of(a.sidx[x.last])];
          append = true;
        } else {
          op = OP_GET;
        }
        a.uop[g] = op;
      }
    }
    const uint64_t want = ballot(append);
    uint32_t wb = 0;
    if (lane_id() == 0 && want) wb = atomicAdd(&s_nins, (uint32_t)__popcll(want));  // LDS
    wb = (uint32_t)__shfl((int)wb, 0, WAVE);
    if (append) s_ins[wb + (uint32_t)__popcll(want & lanemask_lt())] = g;
  }
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
  return (size_t)ntiles * sizeof(Seg);
}

// a.n is the element count, or with a.desc the largest count the scratch was sized for.
hipError_t fold_batch(const FoldArgs& a, hipStream_t stream) {
  const uint32_t ntiles = (a.n + FOLD_TILE - 1) / FOLD_TILE;
  if (a.n == 0) return hipSuccess;
  const uint32_t grid = std::min<uint32_t>(ntiles, 2048);
  Seg* tiles = (Seg*)a.tile_scratch;
  k_fold_reduce<<<grid, FOLD_BLOCK, 0, stream>>>(a, tiles);
  k_fold_scan_tiles<<<1, FOLD_BLOCK, 0, stream>>>(tiles, a, a.counters);
  k_fold_apply<<<grid, FOLD_BLOCK, 0, stream>>>(a, tiles);
  return hipGetLastError();
}

}  // namespace stm
