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
};

// ------------------------------------------------------------------ radix sort (stm_sort.hip)
// Words of scratch (u32) the sort needs for n elements.
size_t sort_hist_words(uint64_t n);
// Stable ascending sort of keys[begin_bit, end_bit) with the original index as payload.
// keys_in is not modified; results land in keys_out / idx_out.
hipError_t radix_sort_pairs(const uint64_t* keys_in, uint64_t* keys_tmp, uint64_t* keys_out,
                            uint32_t* idx_tmp, uint32_t* idx_out, uint32_t n, int begin_bit,
                            int end_bit, uint32_t* hist_words, hipStream_t stream);
// Same with a choice of tile (tile_items = 8 or 16 elements per thread: tiles of 2048 or 4096
// keys).  With desc the keys and their count are read from the device-side descriptor
// (keys_in unused) and n_max only bounds that count (grid and scratch sizing); without it
// keys_in holds n_max keys.
hipError_t radix_sort_pairs_desc(const uint64_t* keys_in, const BatchDesc* desc, uint32_t n_max,
                                 uint64_t* keys_tmp, uint64_t* keys_out, uint32_t* idx_tmp,
                                 uint32_t* idx_out, int begin_bit, int end_bit, uint32_t* hist_words,
                                 int tile_items, hipStream_t stream);

// ------------------------------------------------------------------ per-key fold (stm_fold.hip)
// Per-batch counters written by the fold and the tree kernels (one 64-byte block).
struct BatchCounters {
  uint32_t n_unique;    // groups (unique keys)
  uint32_t n_upserts;   // groups whose merged op is INSERT
  uint32_t n_deferred;  // upserts moved to the serial pass
  uint32_t queue_head;  // work-queue cursor of the upsert kernel
  uint32_t n_bad;       // ops rejected (reserved key)
  uint32_t pad[11];
};

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
};
constexpr uint8_t UOP_SKIP = 0xff;
size_t fold_scratch_bytes(uint64_t n);
hipError_t fold_batch(const FoldArgs& a, hipStream_t stream);

}  // namespace stm
