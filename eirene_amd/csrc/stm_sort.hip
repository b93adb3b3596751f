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
//               1024-element segment, so (wave, round, lane) order is input order), LDS
//               staging into digit order, then coalesced runs to global memory.
// No kernel waits on another workgroup, so no launch order can deadlock.
#include "stm_common.hpp"
#include "stm_internal.hpp"

namespace stm {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;  // 4096
constexpr int SORT_WAVES = SORT_BLOCK / WAVE;       // 4
constexpr int SORT_SEG = SORT_TILE / SORT_WAVES;    // 1024 elements per wavefront
constexpr int RADIX = 256;

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

__global__ __launch_bounds__(SORT_BLOCK) void k_hist(const uint64_t* __restrict__ keys, uint32_t n,
                                                     int shift, uint32_t* __restrict__ hist,
                                                     uint32_t ntiles) {
  __shared__ uint32_t s_h[SORT_WAVES][RADIX];
  const int wave = threadIdx.x / WAVE;
  for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
#pragma unroll 4
  for (int r = 0; r < SORT_ITEMS; ++r) {
    uint64_t e = base + (uint64_t)r * SORT_BLOCK + threadIdx.x;
    if (e < n) {
      uint32_t d = (uint32_t)(keys[e] >> shift) & (RADIX - 1);
      atomicAdd(&s_h[wave][d], 1u);
    }
  }
  __syncthreads();
  const int d = threadIdx.x;  // SORT_BLOCK == RADIX
  uint32_t c = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
  hist[(uint64_t)d * ntiles + blockIdx.x] = c;
}

// Block d scans hist[d][0..ntiles) in place (exclusive) and writes totals[d].
__global__ __launch_bounds__(SORT_BLOCK) void k_scan_hist(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                          uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_tmp[SORT_WAVES];
  uint32_t* row = hist + (uint64_t)blockIdx.x * ntiles;
  const uint32_t per = (ntiles + SORT_BLOCK - 1) / SORT_BLOCK;
  const uint32_t b0 = threadIdx.x * per;
  const uint32_t b1 = min(b0 + per, ntiles);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += row[b];
  uint32_t total;
  uint32_t run = block_exclusive_scan_256(sum, s_tmp, total);
  for (uint32_t b = b0; b < b1; ++b) {
    uint32_t c = row[b];
    row[b] = run;
    run += c;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_scatter(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, uint32_t ntiles, const uint32_t* __restrict__ totals,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out) {
  __shared__ uint32_t s_wcnt[SORT_WAVES][RADIX];  // running per-wave digit counts, then wave bases
  __shared__ uint32_t s_dbase[RADIX];             // first tile position of each digit
  __shared__ uint32_t s_gbase[RADIX];             // global position of each digit's first element
  __shared__ uint32_t s_tmp[SORT_WAVES];
  __shared__ uint64_t s_keys[SORT_TILE];
  __shared__ uint32_t s_idx[SORT_TILE];

  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  const uint64_t tile0 = (uint64_t)blockIdx.x * SORT_TILE;
  const uint32_t tile_n = (uint32_t)min<uint64_t>(SORT_TILE, n - tile0);

  for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_wcnt[0][0])[i] = 0;
  {
    // Global digit bases: exclusive scan of the digit totals.
    uint32_t tot;
    uint32_t tb = block_exclusive_scan_256(totals[threadIdx.x], s_tmp, tot);
    s_gbase[threadIdx.x] = tb + hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  }
  __syncthreads();

  uint64_t k[SORT_ITEMS];
  uint32_t id[SORT_ITEMS];
  uint32_t rk[SORT_ITEMS];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t le = (uint32_t)wave * SORT_SEG + (uint32_t)r * WAVE + lane;  // tile-local
    const bool valid = le < tile_n;
    const uint64_t e = tile0 + le;
    k[r] = valid ? keys_in[e] : KEY_EMPTY;
    id[r] = valid ? (idx_in ? idx_in[e] : (uint32_t)e) : 0u;
    const uint32_t d = (uint32_t)(k[r] >> shift) & (RADIX - 1);
    // Lanes holding the same digit: intersect 8 bit-ballots.
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t base = valid ? s_wcnt[wave][d] : 0u;
    rk[r] = base + (uint32_t)__popcll(peers & lt);
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
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t le = (uint32_t)wave * SORT_SEG + (uint32_t)r * WAVE + lane;
    if (le < tile_n) {
      const uint32_t d = (uint32_t)(k[r] >> shift) & (RADIX - 1);
      const uint32_t p = s_dbase[d] + s_wcnt[wave][d] + rk[r];
      s_keys[p] = k[r];
      s_idx[p] = id[r];
    }
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < tile_n; p += SORT_BLOCK) {
    const uint64_t kk = s_keys[p];
    const uint32_t d = (uint32_t)(kk >> shift) & (RADIX - 1);
    const uint64_t dst = (uint64_t)s_gbase[d] + (p - s_dbase[d]);
    keys_out[dst] = kk;
    idx_out[dst] = s_idx[p];
  }
}

__global__ void k_iota_copy(const uint64_t* __restrict__ keys_in, uint64_t* __restrict__ keys_out,
                            uint32_t* __restrict__ idx_out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys_out[i] = keys_in[i];
    idx_out[i] = i;
  }
}

size_t sort_hist_words(uint64_t n) {
  uint64_t ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  if (ntiles == 0) ntiles = 1;
  return (size_t)ntiles * RADIX + RADIX;
}

hipError_t radix_sort_pairs(const uint64_t* keys_in, uint64_t* keys_tmp, uint64_t* keys_out,
                            uint32_t* idx_tmp, uint32_t* idx_out, uint32_t n, int begin_bit,
                            int end_bit, uint32_t* hist_words, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  uint32_t* hist = hist_words;
  uint32_t* totals = hist_words + (size_t)ntiles * RADIX;
  const int passes = (end_bit - begin_bit + 7) / 8;
  if (passes <= 0) {
    k_iota_copy<<<(n + 255) / 256, 256, 0, stream>>>(keys_in, keys_out, idx_out, n);
    return hipGetLastError();
  }
  const uint64_t* kin = keys_in;
  const uint32_t* iin = nullptr;
  for (int p = 0; p < passes; ++p) {
    const int shift = begin_bit + 8 * p;
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* iout = to_out ? idx_out : idx_tmp;
    k_hist<<<ntiles, SORT_BLOCK, 0, stream>>>(kin, n, shift, hist, ntiles);
    k_scan_hist<<<RADIX, SORT_BLOCK, 0, stream>>>(hist, ntiles, totals);
    k_scatter<<<ntiles, SORT_BLOCK, 0, stream>>>(kin, iin, n, shift, hist, ntiles, totals, kout, iout);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = kout;
    iin = iout;
  }
  return hipSuccess;
}

}  // namespace stm
