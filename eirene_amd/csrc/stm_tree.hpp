// stm_tree.hpp — host launchers of the tree kernels (stm_tree.hip).
#pragma once
#include "stm_common.hpp"
#include "stm_internal.hpp"
#include "stm_learned.hpp"

namespace stm {

// Wavefronts of the upsert fast path at most (2048 workgroups of 4); each records how many ops
// it handed back.
constexpr uint32_t FAST_WAVES_MAX = 8192;

// Answers of a GET group with more than FANOUT_DIRECT waiting ops (a hot key) are not written
// by the one lane that looked the key up: the lookup kernel queues the group's answer range in
// chunks of FANOUT_CHUNK ops and k_fanout writes them with the whole grid.  A batch of n ops
// queues at most fanout_capacity(n) chunks.
constexpr uint32_t FANOUT_DIRECT = 32;
constexpr uint32_t FANOUT_CHUNK = 1024;
inline uint64_t fanout_capacity(uint64_t n) { return n / FANOUT_CHUNK + n / FANOUT_DIRECT + 1; }

// Slot hint of a merged INSERT (written by the lookup kernel): bits 0..7 hold the key's slot
// in the hinted leaf when the key is there (UPOS_FOUND set); bits 16..23 the leaf's entry
// count.  Valid while the leaf keeps the version the lookup kernel read.
constexpr uint32_t UPOS_SLOT = 0xffu;
constexpr uint32_t UPOS_FOUND = 0x100u;
constexpr uint32_t UPOS_COUNT_SHIFT = 16;

// What the lookup kernel records for a merged INSERT (at its group index): everything the
// upsert kernels need of the op apart from its key and value, in one 32-byte record.  The
// inserts are a sparse subset of the groups, so each array they gathered from instead cost a
// line per insert.
struct alignas(32) InsHint {
  uint64_t ver;   // the hinted leaf's lock word when the lookup kernel ran
  uint32_t leaf;  // the leaf whose fences hold the key
  uint32_t pos;   // slot hint (UPOS_*)
  uint32_t par;   // level-1 node that routed it (NODE_NONE if none)
  uint32_t gs;    // sorted position of the group's first op
  uint32_t np;    // ops of the group waiting for the tree's answer
  uint32_t pad;
};
static_assert(sizeof(InsHint) == 32, "two 16-byte words");

struct LookupArgs {
  const uint64_t* ukey;
  const uint8_t* uop;
  const uint32_t* gstart;
  const uint32_t* npend;
  const uint32_t* sidx;
  const BatchCounters* counters;
  uint64_t n_max;     // upper bound of the group count (grid sizing)
  uint32_t chunk;     // sorted groups walked by one lane group
  uint32_t lanes;     // lanes per op: 8, 16, 32, 64
  uint64_t* out_vals;
  uint8_t* out_found;
  // learned routing (all null for root-to-leaf descent)
  const RouteHeader* route_hdr;
  const uint64_t* fences;
  const uint32_t* lids;
  const SegModel* segs;
  uint32_t index;         // learned routing: 1 = search the directory instead of the model
  const BatchDesc* desc;  // optional: out_vals / out_found are taken from here
  InsHint* ihint;         // optional: the record of every merged INSERT (the upsert kernels' hint)
  uint64_t* gval;         // a hot GET group's value (per group; GET groups leave uval unused)
  uint64_t* fanout;       // queued answer chunks: group | first op << 32 | found << 63
  uint64_t fanout_cap;    // fanout_capacity(n_max)
};

struct UpsertArgs {
  const uint64_t* ukey;
  const uint64_t* uval;
  const uint8_t* uop;
  const uint32_t* gstart;
  const uint32_t* npend;
  const uint32_t* sidx;
  BatchCounters* counters;
  uint32_t* dlist;
  uint32_t max_attempts;
  uint32_t blocks;     // 256-thread workgroups; each wavefront owns a range of groups
  uint64_t* out_vals;
  uint8_t* out_found;
  const BatchDesc* desc;
  const InsHint* ihint;   // the lookup kernel's records of the merged INSERTs; with slist: the
                          //   fast path runs
  const uint32_t* ilist;  // the merged INSERT groups (counters->n_upserts)
  uint32_t* slist;        // ops the fast path hands back (capacity: groups), in the handing
                          //   wavefront's own range of the insert list
  uint32_t* scount;       // per fast-path wavefront: ops it handed back (capacity: FAST_WAVES_MAX)
  uint32_t lanes;         // lanes per op of the fast path: 8, 16 or 32
  const uint64_t* fanout; // hot GET groups' answer chunks queued by the lookup kernel (the fast
  uint64_t fanout_cap;    //   path writes them first); null: none queued
  const uint64_t* gval;   // their values (per group)
  BatchCounters* host_counters;  // optional: pinned host memory the last kernel copies the
                                 //   counters into (null: the caller copies them back)
};

hipError_t launch_build_leaves(const TreeView& t, const uint64_t* ukey, const uint64_t* uval,
                               uint32_t U, uint32_t fill, uint32_t L, uint64_t* lows, hipStream_t s);
hipError_t launch_build_inner(const TreeView& t, const uint64_t* lows_prev, uint32_t n_prev,
                              uint32_t first_prev, uint32_t fill, uint32_t n_this,
                              uint32_t first_this, uint32_t level, uint64_t* lows_out,
                              hipStream_t s);
hipError_t launch_lookup(const TreeView& t, const LookupArgs& a, hipStream_t s);
hipError_t launch_upsert(const TreeView& t, const UpsertArgs& a, hipStream_t s);
hipError_t launch_collect_leaves(const TreeView& t, uint32_t n_nodes, uint64_t* lows, uint32_t* ids,
                                 uint32_t* counter, hipStream_t s);
hipError_t launch_leaf_counts(const TreeView& t, const uint32_t* order, const uint32_t* ids,
                              uint32_t n_leaves, uint32_t* sorted_ids, uint32_t* counts,
                              hipStream_t s);
hipError_t launch_dump_copy(const TreeView& t, const uint32_t* sorted_ids, const uint64_t* offsets,
                            uint32_t n_leaves, uint64_t* out_k, uint64_t* out_v, hipStream_t s);
hipError_t launch_check(const TreeView& t, uint32_t n_nodes, unsigned long long* bad, hipStream_t s);

}  // namespace stm
