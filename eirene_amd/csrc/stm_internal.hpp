// stm_internal.hpp — host-side interfaces shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "stm_common.hpp"

namespace stm {

// Device-side description of one batch: where its keys / values / op codes / outputs live and
// how many ops it holds.  Kernels that take one read it at run time, so one captured graph can be
// replayed for any batch of up to the size it was captured for.
struct BatchDesc {
  const uint64_t* keys;
  const uint64_t* vals;
  const uint8_t* ops;
  uint64_t* out_vals;
  uint8_t* out_found;
  uint32_t n;
  uint32_t pad;
  // range-scan output area (scan mode only)
  uint64_t* scan_keys;
  uint64_t* scan_vals;
  uint64_t scan_cap;
};

// Sort payload of a batch element: its input position in bits 0..29 and its op code
// (0..3, codes above 3 stored as GET) in bits 30..31, so the fold reads op codes from the
// sorted payload instead of gathering them.  max_batch <= 2^30 keeps positions in range.
constexpr uint32_t POS_BITS = 30;
constexpr uint32_t POS_MASK = (1u << POS_BITS) - 1;
__host__ __device__ inline uint32_t pos_of(uint32_t payload) { return payload & POS_MASK; }
__host__ __device__ inline uint32_t op_of(uint32_t payload) { return payload >> POS_BITS; }

// ------------------------------------------------------------------ radix sort (stm_sort.hip)
// Words of scratch (u32) the sort needs for n elements.
size_t sort_hist_words(uint64_t n);
// Prepares freshly allocated sort scratch (synchronous): the single-sweep look-back status
// must start with no valid tag.
hipError_t sort_scratch_init(uint32_t* words);
// Non-zero after a sort whose look-back gave up waiting (its output is then wrong); cleared
// when the next sort on the same scratch starts.
const uint32_t* sort_stall_word(const uint32_t* words);
// The three-kernel region of the scratch (at least n/8 words for scratch sized for n), which
// single-sweep sorts leave alone and three-kernel sorts write before they read: free for other
// kernels between sorts.
uint32_t* sort_spare_words(uint32_t* words);
// Stable ascending sort of keys[begin_bit, end_bit) with the original index as payload.
// keys_in is not modified; results land in keys_out / idx_out.
hipError_t radix_sort_pairs(const uint64_t* keys_in, uint64_t* keys_tmp, uint64_t* keys_out,
                            uint32_t* idx_tmp, uint32_t* idx_out, uint32_t n, int begin_bit,
                            int end_bit, uint32_t* hist_words, hipStream_t stream);
// Same with a choice of tile (tile_items = 8 or 16 elements per thread: tiles of 2048 or 4096
// keys) and op codes packed into the payload (pos | op << POS_BITS; ops_in null = all 0).
// With desc the keys, op codes and their count are read from the device-side descriptor
// (keys_in and ops_in unused) and n_max only bounds that count (grid and scratch sizing);
// without it keys_in holds n_max keys.
// Above 2^18 keys the normalized-key form runs unless `exact` asks for the eight-pass form
// (the library asks for it after a batch whose normalized sort had to fall back).
hipError_t radix_sort_pairs_desc(const uint64_t* keys_in, const uint8_t* ops_in,
                                 const BatchDesc* desc, uint32_t n_max,
                                 uint64_t* keys_tmp, uint64_t* keys_out, uint32_t* idx_tmp,
                                 uint32_t* idx_out, int begin_bit, int end_bit, uint32_t* hist_words,
                                 int tile_items, bool exact, hipStream_t stream);
// Non-zero after a normalized-key sort that found a long bucket of distinct keys and re-sorted
// the batch with the eight exact passes inside one launch (correct, but slower); cleared when
// the next sort on the same scratch starts.
const uint32_t* sort_exact_word(const uint32_t* words);

// ------------------------------------------------------------------ per-key fold (stm_fold.hip)
// Per-batch counters written by the fold and the tree kernels (one 64-byte block).
struct BatchCounters {
  uint32_t n_unique;    // groups (unique keys)
  uint32_t n_upserts;   // groups whose merged op is INSERT
  uint32_t n_deferred;  // upserts moved to the serial pass
  uint32_t queue_head;  // work-queue cursor of the upsert kernel
  uint32_t n_bad;       // ops rejected (reserved key)
  uint32_t scan_bad;    // scan ops rejected (length or output area)
  uint64_t scan_total;  // sum of the scan lengths (pairs the output area must hold)
  uint32_t sort_stall;  // the batch's sort gave up a look-back wait
  uint32_t n_fanout;    // answer chunks of hot GET groups queued for k_fanout
  uint64_t hdr_error;   // the tree's sticky error bits after the batch (one readback per batch)
  uint32_t sort_fallback;  // the batch's normalized sort fell back to the exact passes
  uint32_t leaf_splits;    // leaves the batch's transactions split
  uint32_t pad[2];
};
static_assert(sizeof(BatchCounters) == 64, "one 64-byte block");

struct FoldArgs {
  const uint64_t* skeys;  // sorted keys
  const uint32_t* sidx;   // original positions, stable
  const uint64_t* vals;   // original values (may be null if no inserts)
  const uint8_t* ops;     // original op codes, null = every op is an insert
  uint32_t n;             // element count, or the upper bound when desc is set
  const BatchDesc* desc;  // optional: n, vals, ops and outputs come from here
  // outputs
  uint64_t* out_vals;     // per original op (may be null for bulk load)
  uint8_t* out_found;
  uint64_t* ukey;         // per group
  uint64_t* uval;
  uint8_t* uop;           // OP_GET, OP_INSERT or UOP_SKIP
  uint32_t* gstart;       // sorted position of the group's first op
  uint32_t* npend;        // ops answered by the tree (prefix of the group)
  BatchCounters* counters;
  void* tile_scratch;     // fold_scratch_bytes(n)
  int32_t* lastins;       // scan mode: per sorted entry, the latest insert of its key at or
                          // before it (sorted index), -1 if none; null otherwise
  uint32_t* ilist;        // optional: the groups whose merged op is INSERT (n_upserts of them,
                          // key order within each fold tile)
  const uint32_t* sort_stall;  // optional: the sort's stall word, copied into the counters
  const uint32_t* sort_exact;  // optional: the sort's fallback word, copied into the counters
};
constexpr uint8_t UOP_SKIP = 0xff;
constexpr uint8_t OP_RANGE = 3;
constexpr uint32_t SCAN_MAX_LEN = 65536;  // longest scan (pairs)
size_t fold_scratch_bytes(uint64_t n);
hipError_t fold_batch(const FoldArgs& a, hipStream_t stream);

// ------------------------------------------------------------------ range scans (stm_scan.hip)
struct ScanArgs {
  // the batch (or desc)
  const uint64_t* keys;
  const uint64_t* vals;
  const uint8_t* ops;
  uint32_t n;
  const BatchDesc* desc;
  uint64_t* out_vals;
  uint8_t* out_found;
  uint64_t* scan_keys;
  uint64_t* scan_vals;
  uint64_t scan_cap;
  // fold output
  const uint64_t* ukey;
  const uint8_t* uop;
  const uint32_t* gstart;
  const uint32_t* npend;
  const uint32_t* sidx;
  const int32_t* lastins;
  BatchCounters* counters;
  // scratch
  uint64_t* soff;       // per input position: output offset, then pair count
  uint64_t* tile_sums;  // scan_tile_words(n)
};
size_t scan_tile_words(uint64_t n);
hipError_t launch_scan_offsets(const ScanArgs& a, hipStream_t s);  // before the scan kernel
hipError_t launch_range_scan(const TreeView& t, const ScanArgs& a, hipStream_t s);  // pre-batch tree
hipError_t launch_scan_answers(const ScanArgs& a, hipStream_t s);  // after the upsert kernel

}  // namespace stm
