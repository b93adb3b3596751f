// stm_sort.hip — stable LSD radix sort of (u64 key, u32 index) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs<int,int> of the reference
// (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:293, sized at :358) for 64-bit keys.
// Stability is what makes the per-key fold reproduce input order
// (src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:16-44), so every pass is stable.
//
// One pass = 8 key bits.  Full 64-bit keys are sorted by an order-preserving normalized image
// of 16 or 24 bits (k_range, then k_fix orders each bucket by exact key; see the normalized-key
// form below), with the eight exact passes as the fallback.  Two forms of the passes, chosen by
// the largest batch a launch sequence serves:
//
// Up to ONESWEEP_MAX_N keys, single sweep (2 or 3 normalized passes, 8 exact):
//   k_digit_hist  one read of the keys builds the histograms of every digit position
//                 (per-workgroup partials)
//   k_digit_scan  sums the partials and scans them into each pass's global digit offsets
//   k_onesweep    one launch per pass: ranking, the tile's offset among earlier tiles by
//                 decoupled look-back, staging and coalesced writes (as k_scatter below)
// Larger, three kernels per pass (3 normalized passes, 4 above 2^23 keys; 8 exact):
//   k_hist      tile digit histograms           -> hist[digit][tile]
//   k_scan_hist exclusive scan per digit over tiles, digit totals
//   k_scatter   stable in-tile ranking with 64-lane ballots (a wavefront owns a contiguous
//               segment, so (wave, round, lane) order is input order), LDS staging into
//               digit order (keys, then payloads, through one buffer), then coalesced runs
//               to global memory.
// Small batches pay mostly fixed per-kernel costs, which the single sweep cuts by more than
// half.  At 2^22 keys about 1024 tiles of a pass are resident at once and the look-back chain,
// through agent-scope status loads across the XCD caches, costs more per pass (43.6 µs against
// 26.7) than the histogram and scan kernels it replaces.
//
// Look-back: a tile publishes its digit counts as soon as it has ranked (an aggregate), then
// sums the aggregates of the tiles before it until it meets an inclusive prefix, and publishes
// its own inclusive prefix.  Each status is one 64-bit word (tag, inclusive bit, count), so a
// single agent-scope load reads it whole; the tag carries a per-sort epoch and the pass, so
// the status array is never cleared between sorts.
// Progress: workgroups take tiles from per-class tickets (class = blockIdx.x & 7, tile =
// ticket * 8 + class), so within a class tiles are taken in order by running workgroups.  The
// smallest untaken tile then belongs to a class whose running workgroups only wait on smaller,
// already taken tiles, whose aggregates need no waiting: every wait ends.  A wait that still
// exceeds LOOKBACK_SPIN_LIMIT polls sets the stall word and proceeds, so a fault in this
// reasoning reports an error instead of hanging the device.
//
// The element count and the input keys can come from a device-side batch descriptor, and the
// grids are sized for the largest batch, so one launch sequence (and one captured hipGraph)
// serves any batch size up to the one it was sized for.  The tile is SORT_BLOCK * ITEMS
// elements (ITEMS = 8 or 16, chosen by the autotuner).
#include <algorithm>
#include <atomic>

#include "stm_common.hpp"
#include "stm_internal.hpp"

namespace stm {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_WAVES = SORT_BLOCK / WAVE;  // 4
constexpr int RADIX = 256;
constexpr uint32_t SORT_GRID_CAP = 2048;       // workgroups of the tile kernels
constexpr uint32_t SCAN_REG = 8;               // tiles per thread the histogram scan keeps in registers

// Single sweep.
// Largest batch sorted in a single sweep.  Batch of 2^k ops (2^25 keys, 90/10): at k = 17 and 18
// the single sweep took 0.276 and 0.313 ms against 0.313 and 0.327; at 19 and 20 0.412 and
// 0.563 against 0.394 and 0.490.
constexpr uint64_t ONESWEEP_MAX_N = 1u << 18;
constexpr uint32_t OS_MAX_TILES = (uint32_t)(ONESWEEP_MAX_N / (SORT_BLOCK * 8));  // 2048-key tiles
constexpr uint32_t TICKET_CLASSES = 8;         // ticket counters per pass
constexpr int MAX_PASSES = 8;
constexpr int HIST_ITEMS = 8;                  // keys in flight per thread in k_digit_hist
constexpr uint32_t HIST_GRID_CAP = (uint32_t)(ONESWEEP_MAX_N / (SORT_BLOCK * HIST_ITEMS));
constexpr uint32_t LOOKBACK_SPIN_LIMIT = 1u << 18;
constexpr uint32_t EPOCH_LIMIT = 1u << 29;     // epoch << 3 | pass fits the 32-bit tag
constexpr uint64_t STATUS_INCL = 1ull << 31;
constexpr uint32_t STATUS_COUNT = (1u << 31) - 1;

// Sort scratch (u32 words).  First a fixed single-sweep region: control words (ticket counters
// 128 bytes apart), the global digit offsets of every pass, the per-workgroup partial
// histograms and the u64 look-back status of every tile; then the three-kernel region
// (per-tile histograms and digit totals).
constexpr uint32_t CTL_TICKETS = 0;  // [MAX_PASSES][TICKET_CLASSES] * 32
constexpr uint32_t CTL_EPOCH = MAX_PASSES * TICKET_CLASSES * 32;
constexpr uint32_t CTL_WRAPPED = CTL_EPOCH + 32;
constexpr uint32_t CTL_STALL = CTL_WRAPPED + 32;
constexpr uint32_t CTL_NEED_EXACT = CTL_STALL + 32;  // normalized sort: a long bucket holds distinct keys
constexpr uint32_t CTL_BARRIER = CTL_NEED_EXACT + 32;  // grid-barrier arrivals of k_sort_exact
constexpr uint32_t CTL_NORM = CTL_BARRIER + 32;  // u64 kmin, then u32 nshift (first k_hist pass, workgroup 0)
constexpr uint32_t GOFF_OFF = (CTL_NORM + 32 + 63) / 64 * 64;       // goff[MAX_PASSES][RADIX]
constexpr uint32_t PART_OFF = GOFF_OFF + MAX_PASSES * RADIX;        // [HIST_GRID_CAP][MAX_PASSES][RADIX]
constexpr uint32_t STATUS_OFF = PART_OFF + HIST_GRID_CAP * MAX_PASSES * RADIX;  // u64 [OS_MAX_TILES][RADIX]
// k_range: RANGE_BLOCKS workgroups of RANGE_THREADS, one (min, max) partial each (two rounds of
// eight loads per thread at 2^22 keys; one round of sixteen measured the same).  The partials
// are reduced by every workgroup of the first histogram pass: a last-workgroup reduction inside k_range (an agent-scope release and
// an arrival atomic per workgroup, an acquire in the last) took 16 µs per 32 MiB of keys.
constexpr uint32_t RANGE_BLOCKS = 256;
constexpr uint32_t RANGE_THREADS = 1024;
constexpr int RANGE_WAVES = RANGE_THREADS / WAVE;
constexpr uint32_t RANGE_OFF = STATUS_OFF + 2 * OS_MAX_TILES * RADIX;  // u64 [RANGE_BLOCKS][2]
constexpr uint32_t ONESWEEP_WORDS = RANGE_OFF + 4 * RANGE_BLOCKS;
static_assert(STATUS_OFF % 2 == 0 && RANGE_OFF % 2 == 0 && ONESWEEP_WORDS % 64 == 0, "aligned regions");

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

// The valid lanes of the wavefront that hold the same 8-bit digit d as this lane.  Lane M
// differs from this lane at bit b exactly where ballot(bit b) disagrees with this lane's own
// bit b: in ballot(bit b) XOR (all ones if the own bit is set).  OR-ing those over the 8 bits and
// complementing gives the peers.  The own bit is sign-extended into a word by one bit-field
// extract and the masks are kept as 32-bit halves: about 5 VALU per bit, against 9 for
// intersecting ballot(bit) or its complement chosen by a select.
STM_DEV uint64_t digit_peers(uint32_t d, bool valid) {
  uint32_t dlo = 0, dhi = 0;
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const uint64_t bb = ballot((d >> bit) & 1u);
    const uint32_t own = (uint32_t)((int32_t)(d << (31 - bit)) >> 31);
    dlo |= (uint32_t)bb ^ own;
    dhi |= (uint32_t)(bb >> 32) ^ own;
  }
  return ballot(valid) & ~(((uint64_t)dhi << 32) | dlo);
}

constexpr uint32_t FIX_TILE = 2048;  // sorted positions per k_fix step
constexpr uint32_t FIX_SPAN = 64;    // longest bucket k_fix orders itself
constexpr uint32_t FIX_WIN = FIX_TILE + 2 * FIX_SPAN;
constexpr uint32_t BARRIER_SPIN_LIMIT = 1u << 20;  // polls (most sleep ~2k cycles)

STM_DEV uint32_t bitlen64(uint64_t x) { return x ? 64u - (uint32_t)__clzll(x) : 0u; }

struct NormKey {
  uint64_t kmin;
  uint32_t shift;
};

// The normalization the first histogram pass stored for this sort.
STM_DEV NormKey norm_params(const uint32_t* words) {
  NormKey k;
  k.kmin = *reinterpret_cast<const uint64_t*>(words + CTL_NORM);
  k.shift = words[CTL_NORM + 2];
  return k;
}

STM_DEV uint32_t norm_of(uint64_t key, const NormKey& k) {
  return (uint32_t)((key - k.kmin) >> k.shift);
}

// The sort's normalization from k_range's per-workgroup (min, max) partials, reduced by a whole
// SORT_BLOCK workgroup (every workgroup gets the same result).  NB = 8 * npass bits:
// nshift = max(0, bitlen(kmax - kmin) - NB).
STM_DEV NormKey norm_from_partials(const uint32_t* words, int npass) {
  static_assert(RANGE_BLOCKS == SORT_BLOCK, "one partial per thread");
  __shared__ uint64_t s_mm[2][SORT_WAVES];
  const uint64_t* range = reinterpret_cast<const uint64_t*>(words + RANGE_OFF);
  uint64_t mn = range[2 * threadIdx.x], mx = range[2 * threadIdx.x + 1];
#pragma unroll
  for (int off = WAVE / 2; off >= 1; off >>= 1) {
    const uint64_t a = shfl64(mn, lane_id() ^ off), b = shfl64(mx, lane_id() ^ off);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  const int wave = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    s_mm[0][wave] = mn;
    s_mm[1][wave] = mx;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < SORT_WAVES; ++w) {
    mn = s_mm[0][w] < mn ? s_mm[0][w] : mn;
    mx = s_mm[1][w] > mx ? s_mm[1][w] : mx;
  }
  __syncthreads();
  const uint64_t r = mx >= mn ? mx - mn : 0ull;
  const uint32_t bl = bitlen64(r);
  NormKey k;
  k.kmin = mx >= mn ? mn : 0ull;
  k.shift = bl > (uint32_t)(8 * npass) ? bl - (uint32_t)(8 * npass) : 0u;
  return k;
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

// The three kernels of one pass are device functions too: k_sort_exact runs them in turn
// inside one persistent launch.
// norm: the pass sorts normalized keys (key - kmin) >> nshift from k_range's partials, digit
// `shift` of them; without it the key's own digit.
template <int ITEMS>
STM_DEV void hist_body(const uint64_t* __restrict__ keys, const BatchDesc* __restrict__ desc,
                       uint32_t n_host, int first_pass, int shift, uint32_t* __restrict__ hist,
                       uint32_t stride, const uint32_t* __restrict__ norm, int npass) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  __shared__ uint32_t s_h[SORT_WAVES][RADIX];
  const SortRange rg = sort_range(desc, n_host, keys, first_pass);
  uint64_t kmin = 0;
  if (norm) {
    // The first pass reduces k_range's partials itself (no cross-workgroup step inside
    // k_range), and workgroup 0 stores the result for the later kernels of the sort.
    NormKey nk;
    if (first_pass) {
      nk = norm_from_partials(norm, npass);
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t* w = const_cast<uint32_t*>(norm);
        *reinterpret_cast<uint64_t*>(w + CTL_NORM) = nk.kmin;
        w[CTL_NORM + 2] = nk.shift;
      }
    } else {
      nk = norm_params(norm);
    }
    kmin = nk.kmin;
    shift += (int)nk.shift;
  }
  const uint32_t ntiles = (rg.n + TILE - 1) / TILE;
  const int wave = threadIdx.x / WAVE;
  const uint64_t* kp = rg.keys;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)b * TILE;
    // Every key of the thread in one round of loads (the index clamped into the batch, the
    // clamped keys not counted), then the LDS counts.
    uint64_t kk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint64_t e = base + (uint64_t)r * SORT_BLOCK + threadIdx.x;
      kk[r] = gload(kp + (e < rg.n ? e : (uint64_t)rg.n - 1));
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint64_t e = base + (uint64_t)r * SORT_BLOCK + threadIdx.x;
      if (e < rg.n) atomicAdd(&s_h[wave][digit_of(kk[r] - kmin, shift)], 1u);
    }
    __syncthreads();
    const int d = threadIdx.x;  // SORT_BLOCK == RADIX
    hist[(uint64_t)d * stride + b] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
    __syncthreads();
  }
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_hist(const uint64_t* __restrict__ keys,
                                                     const BatchDesc* __restrict__ desc,
                                                     uint32_t n_host, int first_pass, int shift,
                                                     uint32_t* __restrict__ hist, uint32_t stride,
                                                     uint32_t* __restrict__ words, int norm,
                                                     int npass) {
  if (first_pass && !norm && blockIdx.x == 0 && threadIdx.x == 0) {
    words[CTL_STALL] = 0;  // no look-back in this form (k_range opens a normalized sort)
    words[CTL_NEED_EXACT] = 0;
  }
  hist_body<ITEMS>(keys, desc, n_host, first_pass, shift, hist, stride, norm ? words : nullptr,
                   npass);
}

// Row d of hist (the tiles' counts of digit d) is scanned in place (exclusive) and its total
// written to totals[d]; workgroup g takes rows g, g + gridDim.x, ...
template <int ITEMS>
STM_DEV void scan_hist_body(uint32_t* __restrict__ hist, uint32_t stride,
                            const BatchDesc* __restrict__ desc, uint32_t n_host,
                            uint32_t* __restrict__ totals) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  __shared__ uint32_t s_tmp[SORT_WAVES];
  const uint32_t n = desc ? desc->n : n_host;
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  for (uint32_t dg = blockIdx.x; dg < RADIX; dg += gridDim.x) {
  uint32_t* row = hist + (uint64_t)dg * stride;
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
  if (threadIdx.x == 0) totals[dg] = total;
  __syncthreads();
  }  // rows
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_scan_hist(uint32_t* __restrict__ hist, uint32_t stride,
                                                          const BatchDesc* __restrict__ desc,
                                                          uint32_t n_host, uint32_t* __restrict__ totals) {
  scan_hist_body<ITEMS>(hist, stride, desc, n_host, totals);
}

template <int ITEMS>
STM_DEV void scatter_body(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    const uint8_t* __restrict__ ops_in,
    const BatchDesc* __restrict__ desc, uint32_t n_host, int first_pass, int shift,
    const uint32_t* __restrict__ hist, uint32_t stride, const uint32_t* __restrict__ totals,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out,
    const uint32_t* __restrict__ norm, int npass) {
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
  uint64_t kmin = 0;
  if (norm) {
    const NormKey nk = norm_params(norm);
    kmin = nk.kmin;
    shift += (int)nk.shift;
  }
  const uint32_t ntiles = (rg.n + TILE - 1) / TILE;
  const uint64_t* kin = rg.keys;
  const uint8_t* ops = (first_pass && desc) ? desc->ops : ops_in;  // read on the first pass only
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  {
    uint32_t tot;
    s_tbase[threadIdx.x] = block_exclusive_scan_256(totals[threadIdx.x], s_tmp, tot);
  }
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
      const uint32_t d = digit_of(k[r] - kmin, shift);
      dk[r / 4] |= d << (8 * (r % 4));
      const uint64_t peers = digit_peers(d, valid);
      uint32_t base = valid ? s_wcnt[wave][d] : 0u;
      // Peers below this lane (mbcnt: masked popcount against the lane's own position) and in
      // all; the last of them, the one with no peer above, advances the digit's count.
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
      const uint32_t cnt = (uint32_t)__popcll(peers);
      rk[r] = base + below;
      // Materialise the rank here: left to itself the compiler keeps the 64-bit peer mask of
      // every item alive across the barrier below, which costs occupancy.
      asm volatile("" : "+v"(rk[r]));
      if (valid && below + 1u == cnt) s_wcnt[wave][d] = base + cnt;
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
        const uint32_t d = digit_of(kk - kmin, shift);
        dg[r / 4] |= d << (8 * (r % 4));
        // In bounds whenever the histograms are complete; the check keeps a grid barrier of
        // k_sort_exact that gave up waiting from writing outside the output.
        const uint32_t dst = s_gbase[d] + (p - s_dbase[d]);
        if (dst < rg.n) keys_out[dst] = kk;
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
        const uint32_t dst = s_gbase[d] + (p - s_dbase[d]);
        if (dst < rg.n) idx_out[dst] = s_idx[p];
      }
    }
    __syncthreads();
  }
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_scatter(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    const uint8_t* __restrict__ ops_in,
    const BatchDesc* __restrict__ desc, uint32_t n_host, int first_pass, int shift,
    const uint32_t* __restrict__ hist, uint32_t stride, const uint32_t* __restrict__ totals,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out,
    const uint32_t* __restrict__ norm, int npass) {
  scatter_body<ITEMS>(keys_in, idx_in, ops_in, desc, n_host, first_pass, shift, hist, stride,
                      totals, keys_out, idx_out, norm, npass);
}

// Histograms of every digit position in one read of the keys.  Workgroup w counts a
// contiguous share of the keys into LDS and writes its partials to part[w][pass][digit].
// Workgroup 0 also opens the sort: it advances the epoch (wrapping before the tag would
// overflow), clears the stall word and resets the tickets of every pass.
__global__ __launch_bounds__(SORT_BLOCK) void k_digit_hist(const uint64_t* __restrict__ keys,
                                                           const BatchDesc* __restrict__ desc,
                                                           uint32_t n_host, int begin_bit,
                                                           int passes, uint32_t* __restrict__ words,
                                                           int norm) {
  __shared__ uint32_t s_h[SORT_WAVES][MAX_PASSES][RADIX];
  const SortRange rg = sort_range(desc, n_host, keys, 1);
  const int wave = threadIdx.x / WAVE;
  // norm: the digits are those of the normalized key (key - kmin) >> nshift, 8 * passes bits,
  // from k_range's partials (reduced here by every workgroup; workgroup 0 stores the result).
  uint64_t kmin = 0;
  if (norm) {
    const NormKey nk = norm_from_partials(words, passes);
    kmin = nk.kmin;
    begin_bit += (int)nk.shift;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *reinterpret_cast<uint64_t*>(words + CTL_NORM) = nk.kmin;
      words[CTL_NORM + 2] = nk.shift;
    }
  }
  for (int i = threadIdx.x; i < SORT_WAVES * MAX_PASSES * RADIX; i += SORT_BLOCK)
    (&s_h[0][0][0])[i] = 0;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      const uint32_t e = words[CTL_EPOCH] + 1;
      const bool wrap = e >= EPOCH_LIMIT;
      words[CTL_WRAPPED] = wrap ? 1u : 0u;
      words[CTL_EPOCH] = wrap ? 1u : e;
      words[CTL_STALL] = 0;
    }
    if (threadIdx.x < MAX_PASSES * TICKET_CLASSES) words[CTL_TICKETS + threadIdx.x * 32] = 0;
  }
  __syncthreads();
  constexpr uint32_t ROUND = SORT_BLOCK * HIST_ITEMS;
  const uint32_t share = ((rg.n + gridDim.x - 1) / gridDim.x + ROUND - 1) / ROUND * ROUND;
  const uint64_t b0 = (uint64_t)blockIdx.x * share;
  const uint64_t b1 = min<uint64_t>(b0 + share, rg.n);
  const uint64_t* kp = rg.keys;
  for (uint64_t base = b0; base < b1; base += ROUND) {
    uint64_t k[HIST_ITEMS];
#pragma unroll
    for (int r = 0; r < HIST_ITEMS; ++r) {
      const uint64_t e = base + (uint64_t)r * SORT_BLOCK + threadIdx.x;
      k[r] = gload(kp + min(e, b1 - 1));
    }
#pragma unroll
    for (int r = 0; r < HIST_ITEMS; ++r) {
      if (base + (uint64_t)r * SORT_BLOCK + threadIdx.x < b1) {
#pragma unroll
        for (int p = 0; p < MAX_PASSES; ++p)
          if (p < passes) atomicAdd(&s_h[wave][p][digit_of(k[r] - kmin, begin_bit + 8 * p)], 1u);
      }
    }
  }
  __syncthreads();
  uint32_t* part = words + PART_OFF + (uint64_t)blockIdx.x * MAX_PASSES * RADIX;
  const int d = threadIdx.x;  // SORT_BLOCK == RADIX
  for (int p = 0; p < passes; ++p)
    part[p * RADIX + d] = s_h[0][p][d] + s_h[1][p][d] + s_h[2][p][d] + s_h[3][p][d];
}

// Workgroup p sums the partials of pass p over the hist_blocks histogram workgroups and scans
// them into goff[p][digit]: the sorted position of each digit's first element.  After an
// epoch wrap the workgroups also clear the look-back status, so no stale status can carry a
// tag of the new epochs.
__global__ __launch_bounds__(SORT_BLOCK) void k_digit_scan(uint32_t* __restrict__ words,
                                                           uint32_t hist_blocks) {
  __shared__ uint32_t s_tmp[SORT_WAVES];
  const int p = blockIdx.x;
  const int d = threadIdx.x;
  const uint32_t* part = words + PART_OFF + p * RADIX + d;
  uint32_t sum = 0;
#pragma unroll 8
  for (uint32_t b = 0; b < hist_blocks; ++b) sum += part[(uint64_t)b * MAX_PASSES * RADIX];
  uint32_t total;
  words[GOFF_OFF + p * RADIX + d] = block_exclusive_scan_256(sum, s_tmp, total);
  if (words[CTL_WRAPPED]) {
    uint64_t* status = reinterpret_cast<uint64_t*>(words + STATUS_OFF);
    const uint64_t n = (uint64_t)OS_MAX_TILES * RADIX;
    for (uint64_t i = (uint64_t)blockIdx.x * SORT_BLOCK + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * SORT_BLOCK)
      status[i] = 0;
  }
}

// Sum of digit d's counts over the tiles before tile b: aggregates back to the first
// inclusive prefix.  LOOKBACK_WINDOW statuses are read at once, so the walk back over tiles
// that have only published aggregates advances a window per load latency.  A status whose tag
// is not this pass's has not been published yet: the walk resumes there after a short sleep.
constexpr int LOOKBACK_WINDOW = 16;
STM_DEV uint32_t look_back(const uint64_t* status, uint32_t b, uint32_t d, uint32_t tag,
                           uint32_t* stall) {
  uint32_t acc = 0, spins = 0;
  int64_t j = (int64_t)b - 1;  // next tile to account for
  while (j >= 0) {
    uint64_t st[LOOKBACK_WINDOW];
#pragma unroll
    for (int w = 0; w < LOOKBACK_WINDOW; ++w) {
      const int64_t t = j - w;
      st[w] = t >= 0 ? ld_relaxed(status + (uint64_t)t * RADIX + d) : 0ull;
    }
    int used = 0;
    bool done = false, blocked = false;
#pragma unroll
    for (int w = 0; w < LOOKBACK_WINDOW; ++w) {
      if (!done && !blocked) {
        if (j - w < 0) {
          done = true;  // walked past tile 0 (only reached if no inclusive was met)
        } else if ((uint32_t)(st[w] >> 32) != tag) {
          blocked = true;
        } else {
          acc += (uint32_t)st[w] & STATUS_COUNT;
          ++used;
          if (st[w] & STATUS_INCL) done = true;
        }
      }
    }
    if (done) break;
    j -= used;
    if (blocked) {
      if (++spins >= LOOKBACK_SPIN_LIMIT) {
        __hip_atomic_store(stall, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return acc;
}

STM_DEV uint64_t make_status(uint32_t tag, bool inclusive, uint32_t count) {
  return ((uint64_t)tag << 32) | (inclusive ? STATUS_INCL : 0ull) | count;
}

template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    const uint8_t* __restrict__ ops_in,
    const BatchDesc* __restrict__ desc, uint32_t n_host, int first_pass, int pass, int shift,
    uint32_t* __restrict__ words, uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out,
    int norm) {
  constexpr int TILE = SORT_BLOCK * ITEMS;
  constexpr int SEG = TILE / SORT_WAVES;  // contiguous elements owned by one wavefront
  __shared__ uint32_t s_wcnt[SORT_WAVES][RADIX];  // running per-wave digit counts, then wave bases
  __shared__ uint32_t s_dbase[RADIX];             // first tile position of each digit
  __shared__ uint32_t s_gbase[RADIX];             // global position of this tile's run of each digit
  __shared__ uint32_t s_tmp[SORT_WAVES];
  __shared__ uint32_t s_ticket;
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
  uint64_t kmin = 0;  // norm: digits of the normalized key (k_digit_hist stored kmin, nshift)
  if (norm) {
    const NormKey nk = norm_params(words);
    kmin = nk.kmin;
    shift += (int)nk.shift;
  }
  const uint32_t cls = blockIdx.x & (TICKET_CLASSES - 1);
  uint32_t* ticket = words + CTL_TICKETS + (pass * TICKET_CLASSES + cls) * 32;
  uint32_t* stall = words + CTL_STALL;
  uint64_t* status = reinterpret_cast<uint64_t*>(words + STATUS_OFF);
  const uint32_t tag = (words[CTL_EPOCH] << 3) | (uint32_t)pass;
  const uint32_t goff = words[GOFF_OFF + pass * RADIX + threadIdx.x];  // digit threadIdx.x
  for (;;) {
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_wcnt[0][0])[i] = 0;
    if (threadIdx.x == 0) s_ticket = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t b = s_ticket * TICKET_CLASSES + cls;
    if (b >= ntiles) break;
    const uint64_t tile0 = (uint64_t)b * TILE;
    const uint32_t tile_n = (uint32_t)min<uint64_t>(TILE, rg.n - tile0);

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
      const uint32_t d = digit_of(k[r] - kmin, shift);
      dk[r / 4] |= d << (8 * (r % 4));
      const uint64_t peers = digit_peers(d, valid);
      uint32_t base = valid ? s_wcnt[wave][d] : 0u;
      // Peers below this lane (mbcnt: masked popcount against the lane's own position) and in
      // all; the last of them, the one with no peer above, advances the digit's count.
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
      const uint32_t cnt = (uint32_t)__popcll(peers);
      rk[r] = base + below;
      // Materialise the rank here: left to itself the compiler keeps the 64-bit peer mask of
      // every item alive across the barrier below, which costs occupancy.
      asm volatile("" : "+v"(rk[r]));
      if (valid && below + 1u == cnt) s_wcnt[wave][d] = base + cnt;
    }
    __syncthreads();
    uint64_t* row = status + (uint64_t)b * RADIX;
    uint32_t count;
    {
      const int d = threadIdx.x;
      uint32_t c0 = s_wcnt[0][d], c1 = s_wcnt[1][d], c2 = s_wcnt[2][d], c3 = s_wcnt[3][d];
      count = c0 + c1 + c2 + c3;
      // Publish the aggregate first (tile 0's is already its inclusive prefix).
      st_relaxed(row + d, make_status(tag, b == 0, count));
      uint32_t tot;
      s_dbase[d] = block_exclusive_scan_256(count, s_tmp, tot);
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
    // Their loads are in flight during the look-back below.
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
    {
      const int d = threadIdx.x;
      uint32_t before = 0;
      if (b > 0) {
        before = look_back(status, b, (uint32_t)d, tag, stall);
        st_relaxed(row + d, make_status(tag, true, before + count));
      }
      s_gbase[d] = goff + before;
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
        const uint32_t d = digit_of(kk - kmin, shift);
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

// ================================================================ normalized-key form
// Batches above the single-sweep bound sort an order-preserving image of each key, of
// NB = 8 * npass bits (24, or 32 above 2^23 keys):
//   nk = (key - kmin) >> nshift,  nshift = max(0, bitlen(kmax - kmin) - NB)
// with npass stable passes of the three-kernel form (the digits of nk instead of the key's
// own: three or four passes instead of eight).  Equal nk form buckets of keys within 2^nshift
// of each other, in input order.  k_fix then puts every bucket of at most FIX_SPAN elements in
// key order (rank by key, ties by position: stable).  A longer bucket keeps input order, which
// is already the sorted order when all its keys are equal (the ops of a hot key); if one holds
// distinct keys, k_fix raises CTL_NEED_EXACT and k_sort_exact, one persistent launch that
// otherwise returns at once, re-sorts the whole batch with the eight exact 8-bit passes.
// Minimum and maximum of the batch's keys, one partial per workgroup; the first histogram pass
// reduces the partials into the normalization.  Workgroup 0 also opens the sort: the stall
// word, the exact-sort request and the barrier counter start at zero.
__global__ __launch_bounds__(RANGE_THREADS) void k_range(const uint64_t* __restrict__ keys,
                                                         const BatchDesc* __restrict__ desc,
                                                         uint32_t n_host, uint32_t* __restrict__ words) {
  __shared__ uint64_t s_mm[2][RANGE_WAVES];
  const SortRange rg = sort_range(desc, n_host, keys, 1);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    words[CTL_STALL] = 0;
    words[CTL_NEED_EXACT] = 0;
    words[CTL_BARRIER] = 0;
  }
  uint64_t mn = ~0ull, mx = 0ull;
  const uint64_t stride = (uint64_t)gridDim.x * RANGE_THREADS;
  uint64_t i = (uint64_t)blockIdx.x * RANGE_THREADS + threadIdx.x;
  // Eight loads in flight per thread.
  for (; i + 7 * stride < rg.n; i += 8 * stride) {
    uint64_t k[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) k[u] = gload(rg.keys + i + u * stride);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      mn = k[u] < mn ? k[u] : mn;
      mx = k[u] > mx ? k[u] : mx;
    }
  }
  for (; i < rg.n; i += stride) {
    const uint64_t k = gload(rg.keys + i);
    mn = k < mn ? k : mn;
    mx = k > mx ? k : mx;
  }
#pragma unroll
  for (int off = WAVE / 2; off >= 1; off >>= 1) {
    const uint64_t a = shfl64(mn, lane_id() ^ off), b = shfl64(mx, lane_id() ^ off);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  const int wave = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    s_mm[0][wave] = mn;
    s_mm[1][wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < RANGE_WAVES; ++w) {
      mn = s_mm[0][w] < mn ? s_mm[0][w] : mn;
      mx = s_mm[1][w] > mx ? s_mm[1][w] : mx;
    }
    uint64_t* range = reinterpret_cast<uint64_t*>(words + RANGE_OFF);
    range[2 * blockIdx.x] = mn;
    range[2 * blockIdx.x + 1] = mx;
  }
}

// Buckets (runs of equal nk) into key order.  A step covers FIX_TILE sorted positions with
// FIX_SPAN positions of context on each side in LDS.  A bucket of at most FIX_SPAN elements lies
// wholly in the window, and its element i goes to bucket start + its rank by (key, position).
// Elements of longer buckets stay where they are; a longer bucket with two different adjacent
// keys raises CTL_NEED_EXACT.
// Bucket bounds: the 64 lanes of a wavefront hold 64 consecutive positions, so one ballot of
// "first of its bucket" and one of "last of its bucket" give every lane its bucket's start and
// end when they lie among those positions; only a bucket that reaches past them is walked in LDS
// (at most FIX_SPAN steps each way).  Walking every bucket took 40.5 µs per 2^22 keys, the
// ballots 30.6.
// Only keys go through LDS (17 KB per workgroup); a thread keeps the payloads of its own tile
// positions in registers, so eight workgroups fit on a CU, the most its waves allow, and a
// 2^22-key batch is fixed in one round of workgroups.
constexpr uint32_t FIX_ITEMS = FIX_TILE / SORT_BLOCK;
__global__ __launch_bounds__(SORT_BLOCK) void k_fix(const uint64_t* __restrict__ keys_in,
                                                    const uint32_t* __restrict__ idx_in,
                                                    const BatchDesc* __restrict__ desc,
                                                    uint32_t n_host, uint64_t* __restrict__ keys_out,
                                                    uint32_t* __restrict__ idx_out,
                                                    uint32_t* __restrict__ words, int npass) {
  __shared__ uint64_t s_key[FIX_WIN];  // s_key[x]: key of sorted position t0 - FIX_SPAN + x
  const uint32_t n = desc ? desc->n : n_host;
  const NormKey nkp = norm_params(words);
  const uint32_t tid = threadIdx.x;
  bool need = false;
  for (uint32_t t = blockIdx.x; (uint64_t)t * FIX_TILE < n; t += gridDim.x) {
    const uint32_t t0 = t * FIX_TILE, t1 = min(n, t0 + FIX_TILE);
    // Window positions [lo, hi) are in the batch; LDS index x holds position x + t0 - FIX_SPAN.
    const uint32_t lo_x = t0 >= FIX_SPAN ? 0u : FIX_SPAN - t0;
    const uint32_t hi_x = min(n, t1 + FIX_SPAN) - t0 + FIX_SPAN;
    // One round of loads: the thread's tile positions (key and payload; indexes clamped into the
    // tile, the clamped elements not used) and, for 2 * FIX_SPAN threads, one context key.
    uint64_t k[FIX_ITEMS];
    uint32_t pl[FIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < FIX_ITEMS; ++r) {
      const uint32_t i = min(t0 + r * SORT_BLOCK + tid, t1 - 1);
      k[r] = gload(keys_in + i);
      pl[r] = gload(idx_in + i);
    }
    // Context slot: x = tid on the left (positions t0 - FIX_SPAN ..), x = t1 - t0 + tid on the
    // right (positions t1 ..); tid - FIX_SPAN of the right slots.
    const uint32_t cx = tid < FIX_SPAN ? tid : t1 - t0 + tid;
    const bool cvalid = tid < 2 * FIX_SPAN && cx >= lo_x && cx < hi_x;
    uint64_t ck = 0;
    if (cvalid) ck = gload(keys_in + (t0 + cx - FIX_SPAN));
#pragma unroll
    for (uint32_t r = 0; r < FIX_ITEMS; ++r) {
      const uint32_t x = FIX_SPAN + r * SORT_BLOCK + tid;
      if (t0 + r * SORT_BLOCK + tid < t1) s_key[x] = k[r];
    }
    if (cvalid) s_key[cx] = ck;
    __syncthreads();
    // Not unrolled (the walks would be copied eight times): the key comes from LDS and the
    // payload out of the registers by a chain of selects.
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t upto = (2ull << lane) - 1;  // lanes 0..lane
    const uint64_t from = ~0ull << lane;       // lanes lane..63
#pragma unroll 1
    for (uint32_t r = 0; r < FIX_ITEMS; ++r) {
      if (t0 + r * SORT_BLOCK >= t1) break;
      const uint32_t i = t0 + r * SORT_BLOCK + tid;
      const bool valid = i < t1;
      const uint32_t x = FIX_SPAN + r * SORT_BLOCK + tid;
      const bool has_prev = x > lo_x, has_next = x + 1 < hi_x;
      const uint64_t key = s_key[x];
      const uint64_t kp = s_key[has_prev ? x - 1 : x];
      const uint64_t kn = s_key[has_next ? x + 1 : x];
      const uint32_t v = norm_of(key, nkp);
      const bool first = !has_prev || norm_of(kp, nkp) != v;
      const bool last = !has_next || norm_of(kn, nkp) != v;
      const uint64_t fmask = ballot(valid && first) & upto;
      const uint64_t lmask = ballot(valid && last) & from;
      if (!valid) continue;
      uint32_t dst = i;
      if (!(first && last)) {
        const uint32_t cx = x - lane;  // LDS index of the wavefront's first position
        uint32_t l, e;
        if (fmask) {
          l = cx + 63u - (uint32_t)__clzll(fmask);
        } else {
          l = cx;
          while (l > lo_x && x - l < FIX_SPAN && norm_of(s_key[l - 1], nkp) == v) --l;
        }
        if (lmask) {
          e = cx + (uint32_t)__builtin_ctzll(lmask) + 1u;
        } else {
          e = cx + WAVE;
          while (e < hi_x && e - x <= FIX_SPAN && norm_of(s_key[e], nkp) == v) ++e;
        }
        if (x - l >= FIX_SPAN || e - x > FIX_SPAN || e - l > FIX_SPAN) {
          // A long bucket: keep input order; it is sorted already unless its keys differ.
          if (!first && kp != key) need = true;
        } else {
          uint32_t rank = 0;
          for (uint32_t j = l; j < e; ++j) {
            const uint64_t kj = s_key[j];
            rank += (kj < key || (kj == key && j < x)) ? 1u : 0u;
          }
          dst = t0 + l + rank - FIX_SPAN;
        }
      }
      uint32_t pv = pl[0];
#pragma unroll
      for (uint32_t q = 1; q < FIX_ITEMS; ++q) pv = r == q ? pl[q] : pv;
      keys_out[dst] = key;
      idx_out[dst] = pv;
    }
    __syncthreads();
  }
  if (need) words[CTL_NEED_EXACT] = 1;
}

// Grid barrier of k_sort_exact: every workgroup of the (co-resident) grid arrives, then waits
// until all have.  Each wave first waits for its own stores to be acknowledged by the L2; then
// one lane per workgroup releases at agent scope (the XCD's L2 is written back, so the data is
// visible to the other XCDs), arrives, polls with growing sleeps (fewer pollers contending on
// the counter) and acquires (stale lines dropped) before the workgroup goes on.  A wait beyond
// BARRIER_SPIN_LIMIT polls raises the stall word and returns: the batch is then reported failed
// instead of the device hanging.
STM_DEV void grid_barrier(uint32_t* words, uint32_t target) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(words + CTL_BARRIER, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (__hip_atomic_load(words + CTL_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins >= BARRIER_SPIN_LIMIT) {
        __hip_atomic_store(words + CTL_STALL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      if (spins < 8) __builtin_amdgcn_s_sleep(2);
      else if (spins < 64) __builtin_amdgcn_s_sleep(8);
      else __builtin_amdgcn_s_sleep(32);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// The exact sort, run only when k_fix found a long bucket of distinct keys: the eight 8-bit
// passes of the three-kernel form (histograms, scan, scatter), separated by grid barriers
// instead of launches, from the batch's input to keys_out / idx_out.  The grid must fit on the
// device at once (the launcher sizes it from the occupancy the kernel allows).
template <int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_exact(
    const uint64_t* __restrict__ keys_in, const uint8_t* __restrict__ ops_in,
    const BatchDesc* __restrict__ desc, uint32_t n_host, uint64_t* __restrict__ keys_tmp,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_tmp, uint32_t* __restrict__ idx_out,
    uint32_t* __restrict__ words, uint32_t stride) {
  if (__hip_atomic_load(words + CTL_NEED_EXACT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    return;
  uint32_t* hist = words + ONESWEEP_WORDS;
  uint32_t* totals = hist + (size_t)stride * RADIX;
  const uint64_t* kin = keys_in;
  const uint32_t* iin = nullptr;
  uint32_t arrivals = 0;
  for (int p = 0; p < MAX_PASSES; ++p) {
    const bool to_out = ((MAX_PASSES - 1 - p) % 2) == 0;
    uint64_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* iout = to_out ? idx_out : idx_tmp;
    const int first = p == 0 ? 1 : 0;
    hist_body<ITEMS>(kin, desc, n_host, first, 8 * p, hist, stride, nullptr, 0);
    grid_barrier(words, (++arrivals) * gridDim.x);
    scan_hist_body<ITEMS>(hist, stride, desc, n_host, totals);
    grid_barrier(words, (++arrivals) * gridDim.x);
    scatter_body<ITEMS>(kin, iin, first ? ops_in : nullptr, desc, n_host, first, 8 * p, hist,
                        stride, totals, kout, iout, nullptr, 0);
    grid_barrier(words, (++arrivals) * gridDim.x);
    kin = kout;
    iin = iout;
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
  // the three-kernel region, sized for the smallest tile (the most tiles)
  uint64_t ntiles = (n + sort_tile(8) - 1) / sort_tile(8);
  if (ntiles == 0) ntiles = 1;
  return (size_t)ONESWEEP_WORDS + (size_t)ntiles * RADIX + RADIX;
}

hipError_t sort_scratch_init(uint32_t* words) {
  return hipMemset(words, 0, (size_t)ONESWEEP_WORDS * sizeof(uint32_t));
}

const uint32_t* sort_stall_word(const uint32_t* words) { return words + CTL_STALL; }

uint32_t* sort_spare_words(uint32_t* words) { return words + ONESWEEP_WORDS; }

// Workgroups of k_sort_exact<ITEMS> the device holds at once (its grid barriers need them all
// resident), at most SORT_GRID_CAP.
template <int ITEMS>
static hipError_t exact_grid(uint32_t& grid) {
  static std::atomic<uint32_t> resident{0};
  if (resident.load(std::memory_order_relaxed) == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sort_exact<ITEMS>, SORT_BLOCK, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1 || cus < 1) return hipErrorInvalidConfiguration;
    // Two workgroups per CU at most: the phases are memory-bound, and each further workgroup
    // adds a poller to every grid barrier.
    resident.store((uint32_t)std::min<int64_t>((int64_t)cus * std::min(per_cu, 2), SORT_GRID_CAP),
                   std::memory_order_relaxed);
  }
  grid = resident.load(std::memory_order_relaxed);
  return hipSuccess;
}

// Batches above the single-sweep bound: k_range, npass normalized passes of the three-kernel
// form, k_fix, k_sort_exact (returns at once unless k_fix asked for it).  The passes ping-pong
// between the tmp and out arrays; the last lands in tmp so k_fix can write out.
template <int ITEMS>
static hipError_t sort_normalized(const uint64_t* keys_in, const uint8_t* ops_in,
                                  const BatchDesc* desc, uint32_t n_host, uint32_t n_max,
                                  uint64_t* keys_tmp, uint64_t* keys_out, uint32_t* idx_tmp,
                                  uint32_t* idx_out, uint32_t* words, hipStream_t stream) {
  const uint32_t tile = sort_tile(ITEMS);
  const uint32_t stride = (n_max + tile - 1) / tile;
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(stride, SORT_GRID_CAP));
  const int npass = n_max > (1u << 23) ? 4 : 3;
  uint32_t* hist = words + ONESWEEP_WORDS;
  uint32_t* totals = hist + (size_t)stride * RADIX;
  uint32_t xgrid = 0;
  hipError_t e = exact_grid<ITEMS>(xgrid);
  if (e != hipSuccess) return e;
  k_range<<<RANGE_BLOCKS, RANGE_THREADS, 0, stream>>>(keys_in, desc, n_host, words);
  const uint64_t* kin = keys_in;
  const uint32_t* iin = nullptr;
  for (int p = 0; p < npass; ++p) {
    const bool to_tmp = ((npass - 1 - p) % 2) == 0;
    uint64_t* kout = to_tmp ? keys_tmp : keys_out;
    uint32_t* iout = to_tmp ? idx_tmp : idx_out;
    const int first = p == 0 ? 1 : 0;
    k_hist<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, desc, n_host, first, 8 * p, hist, stride,
                                                   words, 1, npass);
    k_scan_hist<ITEMS><<<RADIX, SORT_BLOCK, 0, stream>>>(hist, stride, desc, n_host, totals);
    k_scatter<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, iin, first ? ops_in : nullptr, desc,
                                                      n_host, first, 8 * p, hist, stride, totals,
                                                      kout, iout, words, npass);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = kout;
    iin = iout;
  }
  const uint32_t fgrid = std::max<uint32_t>(1, std::min<uint32_t>((n_max + FIX_TILE - 1) / FIX_TILE,
                                                                  SORT_GRID_CAP));
  k_fix<<<fgrid, SORT_BLOCK, 0, stream>>>(keys_tmp, idx_tmp, desc, n_host, keys_out, idx_out, words,
                                          npass);
  k_sort_exact<ITEMS><<<xgrid, SORT_BLOCK, 0, stream>>>(keys_in, ops_in, desc, n_host, keys_tmp,
                                                       keys_out, idx_tmp, idx_out, words, stride);
  return hipGetLastError();
}

template <int ITEMS>
static hipError_t sort_passes(const uint64_t* keys_in, const uint8_t* ops_in, const BatchDesc* desc,
                              uint32_t n_host,
                              uint32_t n_max, uint64_t* keys_tmp, uint64_t* keys_out,
                              uint32_t* idx_tmp, uint32_t* idx_out, int begin_bit, int end_bit,
                              uint32_t* words, bool exact, hipStream_t stream) {
  const uint32_t tile = sort_tile(ITEMS);
  const uint32_t stride = (n_max + tile - 1) / tile;  // tiles of the largest batch
  const int passes = (end_bit - begin_bit + 7) / 8;
  if (passes <= 0) {
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(stride, SORT_GRID_CAP));
    k_iota_copy<<<grid, 256, 0, stream>>>(keys_in, desc, n_host, keys_out, idx_out);
    return hipGetLastError();
  }
  if (passes > MAX_PASSES) return hipErrorInvalidValue;
  const uint64_t* kin = keys_in;
  const uint32_t* iin = nullptr;
  if (n_max <= ONESWEEP_MAX_N) {
    // Single sweep.  Full 64-bit keys outside the exact form sort their normalized image
    // (k_range, then the first kernel reduces its partials) in two passes up to 2^14 keys and
    // three above, then k_fix orders each bucket by exact key and k_sort_exact takes over when a
    // long bucket holds distinct keys, as in sort_normalized; the passes ping-pong so that the
    // last one lands in tmp.  (Eight exact passes took 120 µs per 2^17-op batch.)
    const bool normalized = !exact && begin_bit == 0 && end_bit == 64;
    const int npass = normalized ? (n_max <= (1u << 14) ? 2 : 3) : passes;
    uint32_t xgrid = 0;
    if (normalized) {
      hipError_t e = exact_grid<ITEMS>(xgrid);
      if (e != hipSuccess) return e;
      k_range<<<RANGE_BLOCKS, RANGE_THREADS, 0, stream>>>(keys_in, desc, n_host, words);
    }
    // Every ticket class needs workgroups: the grid is a multiple of TICKET_CLASSES.
    const uint32_t grid = (std::max<uint32_t>(stride, 1) + TICKET_CLASSES - 1) / TICKET_CLASSES *
                          TICKET_CLASSES;
    const uint32_t hist_grid = std::max<uint32_t>(
        1, (n_max + SORT_BLOCK * HIST_ITEMS - 1) / (SORT_BLOCK * HIST_ITEMS));
    k_digit_hist<<<hist_grid, SORT_BLOCK, 0, stream>>>(keys_in, desc, n_host, begin_bit, npass,
                                                       words, normalized ? 1 : 0);
    k_digit_scan<<<npass, SORT_BLOCK, 0, stream>>>(words, hist_grid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int p = 0; p < npass; ++p) {
      const int shift = begin_bit + 8 * p;
      const bool last_to_out = ((npass - 1 - p) % 2) == 0;
      const bool to_out = normalized ? !last_to_out : last_to_out;
      uint64_t* kout = to_out ? keys_out : keys_tmp;
      uint32_t* iout = to_out ? idx_out : idx_tmp;
      const int first = p == 0 ? 1 : 0;
      k_onesweep<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, iin, first ? ops_in : nullptr, desc,
                                                         n_host, first, p, shift, words, kout, iout,
                                                         normalized ? 1 : 0);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      kin = kout;
      iin = iout;
    }
    if (normalized) {
      const uint32_t fgrid = std::max<uint32_t>(
          1, std::min<uint32_t>((n_max + FIX_TILE - 1) / FIX_TILE, SORT_GRID_CAP));
      k_fix<<<fgrid, SORT_BLOCK, 0, stream>>>(keys_tmp, idx_tmp, desc, n_host, keys_out, idx_out,
                                              words, npass);
      k_sort_exact<ITEMS><<<xgrid, SORT_BLOCK, 0, stream>>>(keys_in, ops_in, desc, n_host,
                                                           keys_tmp, keys_out, idx_tmp, idx_out,
                                                           words, stride);
      e = hipGetLastError();
    }
    return e;
  }
  if (!exact && begin_bit == 0 && end_bit == 64)
    return sort_normalized<ITEMS>(keys_in, ops_in, desc, n_host, n_max, keys_tmp, keys_out,
                                  idx_tmp, idx_out, words, stream);
  // The exact form: eight 8-bit passes of three kernels each.
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(stride, SORT_GRID_CAP));
  uint32_t* hist = words + ONESWEEP_WORDS;
  uint32_t* totals = hist + (size_t)stride * RADIX;
  for (int p = 0; p < passes; ++p) {
    const int shift = begin_bit + 8 * p;
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* iout = to_out ? idx_out : idx_tmp;
    const int first = p == 0 ? 1 : 0;
    k_hist<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, desc, n_host, first, shift, hist, stride,
                                                   words, 0, 0);
    k_scan_hist<ITEMS><<<RADIX, SORT_BLOCK, 0, stream>>>(hist, stride, desc, n_host, totals);
    k_scatter<ITEMS><<<grid, SORT_BLOCK, 0, stream>>>(kin, iin, first ? ops_in : nullptr, desc,
                                                      n_host, first, shift, hist, stride,
                                                      totals, kout, iout, nullptr, 0);
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
                                 int tile_items, bool exact, hipStream_t stream) {
  if (n_max == 0) return hipSuccess;
  const uint32_t n_host = desc ? 0u : n_max;
  if (tile_items == 8)
    return sort_passes<8>(keys_in, ops_in, desc, n_host, n_max, keys_tmp, keys_out, idx_tmp,
                          idx_out, begin_bit, end_bit, hist_words, exact, stream);
  return sort_passes<16>(keys_in, ops_in, desc, n_host, n_max, keys_tmp, keys_out, idx_tmp,
                         idx_out, begin_bit, end_bit, hist_words, exact, stream);
}

hipError_t radix_sort_pairs(const uint64_t* keys_in, uint64_t* keys_tmp, uint64_t* keys_out,
                            uint32_t* idx_tmp, uint32_t* idx_out, uint32_t n, int begin_bit,
                            int end_bit, uint32_t* hist_words, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  return sort_passes<16>(keys_in, nullptr, nullptr, n, n, keys_tmp, keys_out, idx_tmp, idx_out,
                         begin_bit, end_bit, hist_words, false, stream);
}

const uint32_t* sort_exact_word(const uint32_t* words) { return words + CTL_NEED_EXACT; }

}  // namespace stm
