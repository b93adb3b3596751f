// stm_sort.hip — stable LSD radix sort of (u64 key, u32 index) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs<int,int> of the reference
// (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:293, sized at :358) for 64-bit keys.
// Stability is what makes the per-key fold reproduce input order
// (src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:16-44), so every pass is stable.
//
// One pass = 8 key bits, three kernels:
//   k_hist      tile digit histograms           -> hist[digit][tile]
//   k_scan_hist exclusive scan per digit over tiles, digit totals
//   k_scatter   stable in-tile ranking with 64-lane ballots (a wavefront owns a contiguous
//               segment, so (wave, round, lane) order is input order), LDS staging into
//               digit order (keys, then payloads, through one buffer), then coalesced runs
//               to global memory.
// The element count and the input keys can come from a device-side batch descriptor, and
// the tile kernels loop over tiles with a fixed grid, so one launch sequence (and one
// captured hipGraph) serves any batch size up to the one the scratch was sized for.  The
// tile is SORT_BLOCK * ITEMS elements (ITEMS = 8 or 16, chosen by the autotuner).
// No kernel waits on another workgroup, so no launch order can deadlock.
#include <algorithm>

#include "stm_common.hpp"
#include "stm_internal.hpp"

namespace stm {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_WAVES = SORT_BLOCK / WAVE;  // 4
constexpr int RADIX = 256;
constexpr uint32_t SORT_GRID_CAP = 2048;       // workgroups of the tile kernels
constexpr uint32_t SCAN_REG = 8;               // tiles per thread the histogram scan keeps in registers

// Exclusive scan of one u32 per thread over a 256-thread block; returns the block total.
STM_DEV uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* s_tmp /*[4]*/, uint32_t& total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint32_t t = __shfl_up(x, off, WAVE);
    if (lane >= off) x += t;
  }
  if (lane == WAVE - 1) s_tmp[wave] = x;
  __syncthreads();
  uint32_t wbase = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < SORT_WAVES; ++w) {
    uint32_t s = s_tmp[w];
    if (w < wave) wbase += s;
    total += s;
  }
  __syncthreads();
  return wbase + x - v;
}

// Digit (8 bits from bit `shift`, uniform) of a key with 32-bit operations only.
STM_DEV uint32_t digit_of(uint64_t k, int shift) {
  const uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
  const uint32_t w = shift < 32 ? __builtin_amdgcn_alignbit(hi, lo, (uint32_t)shift)
                                : hi >> (shift - 32);
  return w & (RADIX - 1);
}

struct SortRange {
  uint32_t n;
  const uint64_t* keys;
};

// The descriptor's keys are the input of the first pass only; later passes read the ping-pong
// buffers.
STM_DEV SortRange sort_range(const BatchDesc* d, uint32_t n_host, const uint64_t* keys,
                             int first_pass) {
  SortRange r;
  r.n = d ? d->n : n_host;
  r.keys = (d && first_pass) ? d->keys : keys;
  return r;
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_hist(const uint64_t* __restrict__ keys,
                                                     const BatchDesc* __restrict__ desc,
                                                     uint32_t n_host, int first_pass, int shift,
                                                     uint32_t* __restrict__ hist, uint32_t stride) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  __shared__ uint32_t s_h[SORT_WAVES][RADIX];
  const SortRange rg = sort_range(desc, n_host, keys, first_pass);
  const uint32_t ntiles = (rg.n + TILE - 1) / TILE;
  const int wave = threadIdx.x / WAVE;
  const uint64_t* kp = rg.keys;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)b * TILE;
#pragma unroll 4
    for (int r = 0; r < ITEMS; ++r) {
      const uint64_t e = base + (uint64_t)r * SORT_BLOCK + threadIdx.x;
      if (e < rg.n) atomicAdd(&s_h[wave][digit_of(gload(kp + e), shift)], 1u);
    }
    __syncthreads();
    const int d = threadIdx.x;  // SORT_BLOCK == RADIX
    hist[(uint64_t)d * stride + b] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
    __syncthreads();
  }
}

// Block d scans hist[d][0..ntiles) in place (exclusive) and writes totals[d].
template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_scan_hist(uint32_t* __restrict__ hist, uint32_t stride,
                                                          const BatchDesc* __restrict__ desc,
                                                          uint32_t n_host, uint32_t* __restrict__ totals) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  __shared__ uint32_t s_tmp[SORT_WAVES];
  const uint32_t n = desc ? desc->n : n_host;
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  uint32_t* row = hist + (uint64_t)blockIdx.x * stride;
  const uint32_t per = (ntiles + SORT_BLOCK - 1) / SORT_BLOCK;
  const uint32_t b0 = min(threadIdx.x * per, ntiles);
  const uint32_t b1 = min(b0 + per, ntiles);
  uint32_t total;
  if (per <= SCAN_REG) {
    // Up to SCAN_REG tiles per thread held in registers: one read and one write of the row.
    uint32_t c[SCAN_REG];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_REG; ++j) {
      c[j] = b0 + j < b1 ? row[b0 + j] : 0u;
      sum += c[j];
    }
    uint32_t run = block_exclusive_scan_256(sum, s_tmp, total);
#pragma unroll
    for (int j = 0; j < SCAN_REG; ++j) {
      if (b0 + j < b1) row[b0 + j] = run;
      run += c[j];
    }
  } else {
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += row[b];
    uint32_t run = block_exclusive_scan_256(sum, s_tmp, total);
    for (uint32_t b = b0; b < b1; ++b) {
      uint32_t c = row[b];
      row[b] = run;
      run += c;
    }
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_scatter(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    const uint8_t* __restrict__ ops_in,
    const BatchDesc* __restrict__ desc, uint32_t n_host, int first_pass, int shift,
    const uint32_t* __restrict__ hist, uint32_t stride, const uint32_t* __restrict__ totals,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  constexpr int SEG = TILE / SORT_WAVES;  // contiguous elements owned by one wavefront
  __shared__ uint32_t s_wcnt[SORT_WAVES][RADIX];  // running per-wave digit counts, then wave bases
  __shared__ uint32_t s_dbase[RADIX];             // first tile position of each digit
  __shared__ uint32_t s_tbase[RADIX];             // global position of each digit's first element
  __shared__ uint32_t s_gbase[RADIX];             // global position of this tile's run of each digit
  __shared__ uint32_t s_tmp[SORT_WAVES];
  // Keys, then payloads, staged in digit order through one buffer: 8 bytes per element of
  // LDS instead of 12 lets four 4096-element tiles share a CU.
  __shared__ uint64_t s_stage[TILE];
  uint32_t* s_idx = reinterpret_cast<uint32_t*>(s_stage);

  const SortRange rg = sort_range(desc, n_host, keys_in, first_pass);
  const uint32_t ntiles = (rg.n + TILE - 1) / TILE;
  const uint64_t* kin = rg.keys;
  const uint8_t* ops = (first_pass && desc) ? desc->ops : ops_in;  // read on the first pass only
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  {
    uint32_t tot;
    s_tbase[threadIdx.x] = block_exclusive_scan_256(totals[threadIdx.x], s_tmp, tot);
  }
  const uint64_t lt = lanemask_lt();
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const uint64_t tile0 = (uint64_t)b * TILE;
    const uint32_t tile_n = (uint32_t)min<uint64_t>(TILE, rg.n - tile0);
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_wcnt[0][0])[i] = 0;
    s_gbase[threadIdx.x] = s_tbase[threadIdx.x] + hist[(uint64_t)threadIdx.x * stride + b];
    __syncthreads();

    // Registers bound the tile: keys and ranks live through the ranking, payloads are loaded
    // only once the keys are staged.
    uint64_t k[ITEMS];
    uint32_t rk[ITEMS];
    // Per-tile copies of the lane's indices: opaque to the compiler, so it recomputes the
    // per-item offsets in each tile instead of hoisting 32 of them into registers.
    uint32_t lane0 = (uint32_t)wave * SEG + lane;  // tile-local index of item 0
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(lane0), "+v"(tid));
    // Item r of a lane is tile element lane0 + r * WAVE.  Loads of a partial tile clamp the
    // index into the tile instead of branching; the clamped values are never used.
    const uint32_t last = tile_n - 1;
    const uint64_t* kt = kin + tile0;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t le = lane0 + (uint32_t)r * WAVE;
      const uint64_t v = gload(kt + min(le, last));
      k[r] = le < tile_n ? v : KEY_EMPTY;
    }
    asm volatile("" : "+v"(lane0));  // each phase recomputes its offsets
    uint32_t dk[(ITEMS + 3) / 4];  // the items' digits, four to a register
#pragma unroll
    for (int q = 0; q < (ITEMS + 3) / 4; ++q) dk[q] = 0;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const bool valid = lane0 + (uint32_t)r * WAVE < tile_n;
      const uint32_t d = digit_of(k[r], shift);
      dk[r / 4] |= d << (8 * (r % 4));
      // Lanes holding the same digit: intersect 8 bit-ballots.
      uint64_t peers = ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const bool on = (d >> bit) & 1u;
        const uint64_t bb = ballot(on);
        peers &= on ? bb : ~bb;
      }
      uint32_t base = valid ? s_wcnt[wave][d] : 0u;
      rk[r] = base + (uint32_t)__popcll(peers & lt);
      // Materialise the rank here: left to itself the compiler keeps the 64-bit peer mask of
      // every item alive across the barrier below, which costs occupancy.
      asm volatile("" : "+v"(rk[r]));
      const bool leader = valid && (lane == 63 - __clzll(peers));
      if (leader) s_wcnt[wave][d] = base + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
      const int d = threadIdx.x;
      uint32_t c0 = s_wcnt[0][d], c1 = s_wcnt[1][d], c2 = s_wcnt[2][d], c3 = s_wcnt[3][d];
      uint32_t tot;
      uint32_t db = block_exclusive_scan_256(c0 + c1 + c2 + c3, s_tmp, tot);
      s_dbase[d] = db;
      s_wcnt[0][d] = 0;
      s_wcnt[1][d] = c0;
      s_wcnt[2][d] = c0 + c1;
      s_wcnt[3][d] = c0 + c1 + c2;
    }
    __syncthreads();
    asm volatile("" : "+v"(lane0));
    // rk becomes the element's position in the tile's digit order
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      if (lane0 + (uint32_t)r * WAVE < tile_n) {
        const uint32_t d = (dk[r / 4] >> (8 * (r % 4))) & (RADIX - 1);
        rk[r] += s_dbase[d] + s_wcnt[wave][d];
        s_stage[rk[r]] = k[r];
      }
    }
    asm volatile("" : "+v"(lane0));
    // Payloads: the input's, or on the first pass the position with the op code above it.
    uint32_t id[ITEMS];
    if (idx_in) {
      const uint32_t* it = idx_in + tile0;
#pragma unroll
      for (int r = 0; r < ITEMS; ++r) id[r] = gload(it + min(lane0 + (uint32_t)r * WAVE, last));
    } else if (ops) {
      const uint8_t* ot = ops + tile0;
#pragma unroll
      for (int r = 0; r < ITEMS; ++r) {
        const uint32_t le = lane0 + (uint32_t)r * WAVE;
        const uint32_t o = gload(ot + min(le, last));
        id[r] = ((uint32_t)tile0 + le) | ((o > 3u ? 1u : o) << POS_BITS);
      }
    } else {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r) id[r] = (uint32_t)tile0 + lane0 + (uint32_t)r * WAVE;
    }
    __syncthreads();
    // The digit of every position a thread writes, four to a register, for the payload phase.
    uint32_t dg[(ITEMS + 3) / 4];
#pragma unroll
    for (int q = 0; q < (ITEMS + 3) / 4; ++q) dg[q] = 0;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t p = (uint32_t)r * SORT_BLOCK + tid;
      if (p < tile_n) {
        const uint64_t kk = s_stage[p];
        const uint32_t d = digit_of(kk, shift);
        dg[r / 4] |= d << (8 * (r % 4));
        keys_out[s_gbase[d] + (p - s_dbase[d])] = kk;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      if (lane0 + (uint32_t)r * WAVE < tile_n) s_idx[rk[r]] = id[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t p = (uint32_t)r * SORT_BLOCK + tid;
      if (p < tile_n) {
        const uint32_t d = (dg[r / 4] >> (8 * (r % 4))) & (RADIX - 1);
        idx_out[s_gbase[d] + (p - s_dbase[d])] = s_idx[p];
      }
    }
    __syncthreads();
  }
}

__global__ void k_iota_copy(const uint64_t* __restrict__ keys_in, const BatchDesc* __restrict__ desc,
                            uint32_t n_host, uint64_t* __restrict__ keys_out,
                            uint32_t* __restrict__ idx_out) {
  const SortRange rg = sort_range(desc, n_host, keys_in, 1);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rg.n; i += gridDim.x * blockDim.x) {
    keys_out[i] = rg.keys[i];
    idx_out[i] = i;
  }
}

static uint32_t sort_tile(int items) { return (uint32_t)(SORT_BLOCK * items); }

size_t sort_hist_words(uint64_t n) {
  // sized for the smallest tile (the most tiles)
  uint64_t ntiles = (n + sort_tile(8) - 1) / sort_tile(8);
  if (ntiles == 0) ntiles = 1;
  return (size_t)ntiles * RADIX + RADIX;
}

template <int ITEMS>
static hipError_t sort_passes(const uint64_t* keys_in, const uint8_t* ops_in, const BatchDesc* desc,
                              uint32_t n_host,
                              uint32_t n_max, uint64_t* keys_tmp, uint64_t* keys_out,
                              uint32_t* idx_tmp, uint32_t* idx_out, int begin_bit, int end_bit,
                              uint32_t* hist_words, hipStream_t stream) {
  const uint32_t tile = sort_tile(ITEMS);
  const uint32_t stride = (n_max + tile - 1) / tile;  // tiles of the largest batch
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(stride, SORT_GRID_CAP));
  uint32_t* hist = hist_words;
  uint32_t* totals = hist_words + (size_t)stride * RADIX;
  const int passes = (end_bit - begin_bit + 7) / 8;
  if (passes <= 0) {
    k_iota_copy<<<grid, 256, 0, stream>>>(keys_in, desc, n_host, keys_out, idx_out);
    return hipGetLastError();
  }
  const uint64_t* kin = keys_in;
  const uint32_t* iin = nullptr;
  for (int p = 0; p < passes; ++p) {
    const int shift = begin_bit + 8 * p;
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* iout = to_out ? idx_out : idx_tmp;
    const int first = p == 0 ? 1 : 0;
    k_hist<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, desc, n_host, first, shift, hist, stride);
    k_scan_hist<ITEMS><<<RADIX, SORT_BLOCK, 0, stream>>>(hist, stride, desc, n_host, totals);
    k_scatter<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, iin, first ? ops_in : nullptr, desc,
                                                      n_host, first, shift, hist, stride,
                                                      totals, kout, iout);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = kout;
    iin = iout;
  }
  return hipSuccess;
}

hipError_t radix_sort_pairs_desc(const uint64_t* keys_in, const uint8_t* ops_in,
                                 const BatchDesc* desc, uint32_t n_max,
                                 uint64_t* keys_tmp, uint64_t* keys_out, uint32_t* idx_tmp,
                                 uint32_t* idx_out, int begin_bit, int end_bit, uint32_t* hist_words,
                                 int tile_items, hipStream_t stream) {
  if (n_max == 0) return hipSuccess;
  const uint32_t n_host = desc ? 0u : n_max;
  if (tile_items == 8)
    return sort_passes<8>(keys_in, ops_in, desc, n_host, n_max, keys_tmp, keys_out, idx_tmp,
                          idx_out, begin_bit, end_bit, hist_words, stream);
  return sort_passes<16>(keys_in, ops_in, desc, n_host, n_max, keys_tmp, keys_out, idx_tmp,
                         idx_out, begin_bit, end_bit, hist_words, stream);
}

hipError_t radix_sort_pairs(const uint64_t* keys_in, uint64_t* keys_tmp, uint64_t* keys_out,
                            uint32_t* idx_tmp, uint32_t* idx_out, uint32_t n, int begin_bit,
                            int end_bit, uint32_t* hist_words, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  return sort_passes<16>(keys_in, nullptr, nullptr, n, n, keys_tmp, keys_out, idx_tmp, idx_out,
                         begin_bit, end_bit, hist_words, stream);
}

}  // namespace stm
