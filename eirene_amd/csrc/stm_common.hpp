// stm_common.hpp — HBM layout of the tree and CDNA4 wavefront helpers shared by all kernels.
//
// Node pool (structure of arrays, node id = index):
//   keys[id * F + s]  u64 slot keys, sorted, empty slots hold KEY_EMPTY
//   vals[id * F + s]  u64 leaf values, or child node ids in inner nodes
//   meta[id]          64 B: version/lock word, count|level, low/high fences, right sibling
// Inner nodes use the min-key convention: keys[s] is the low fence of child vals[s], so
// keys[0] equals the node's own low fence and the child for key k is the last slot with
// keys[s] <= k.  Leaves hold (keys[s], vals[s]) pairs.  Every node covers [low, high); the
// high fence of a node equals the low fence of its right sibling (KEY_EMPTY = +inf for
// the rightmost node of a level).  The reference keeps the node header in key slot 31 and
// the leaf's right fence in value slot 31 (src/gpu-stm-tree-v38/bpt-gpustm-kernel-tree.cuh:25-56,
// 139-141); here they live in the separate 64-byte meta record so every key slot is usable.
//
// Lock word (reference: 32-bit version_t with bit 0 = lock, src/gpu-stm-tree-v38/bpt-gpustm-kernel-tx.cuh:16-18):
//   64-bit, bit 0 = locked, bits 63..1 = version.  Unlock stores word + 1 (version + 1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define STM_DEV __device__ __forceinline__

namespace stm {

constexpr uint64_t KEY_EMPTY = ~0ull;
constexpr uint32_t NODE_NONE = 0xffffffffu;
constexpr uint8_t OP_INSERT = 0;
constexpr uint8_t OP_GET = 1;
constexpr int WAVE = 64;

struct alignas(64) NodeMeta {
  uint64_t lock;   // bit 0 locked, version above
  uint64_t info;   // count (bits 0..31) | level (bits 32..63); level 0 = leaf
  uint64_t low;    // inclusive low fence
  uint64_t high;   // exclusive high fence, KEY_EMPTY = +inf
  uint64_t next;   // right sibling id at the same level, NODE_NONE at the right edge
  uint64_t pad[3];
};
static_assert(sizeof(NodeMeta) == 64, "meta record must be one 64-byte sector");

struct alignas(64) TreeHeader {
  uint64_t lock;      // version/lock word guarding root and height
  uint64_t root;      // root node id
  uint64_t height;    // number of levels, 1 = the root is a leaf
  uint64_t n_nodes;   // bump allocator
  uint64_t capacity;  // node-pool capacity
  uint64_t error;     // sticky error bits (ERR_*) raised by kernels
  uint64_t n_keys;    // keys of the last bulk build; inserts since then are counted in ST_KEYS
  uint64_t pad;
};

// Sticky error bits in TreeHeader::error.
constexpr uint64_t ERR_POOL_FULL = 1;     // an insert needed a node beyond capacity
constexpr uint64_t ERR_LIVELOCK = 2;      // the serial pass could not commit an op
constexpr uint64_t ERR_BAD_KEY = 4;       // a batch carried the reserved key
constexpr uint64_t ERR_BAD_OP = 8;        // a batch carried an op code > 3

// Device counters accumulated by kernels.  Each counter is striped over STAT_STRIPES words
// (a wavefront adds to stripe wave_id mod STAT_STRIPES) and the host sums the stripes: one hot
// address per counter would serialize a kernel of thousands of wavefronts on one L2 channel.
enum StatSlot : int {
  ST_COMMITS = 0,
  ST_ABORTS,
  ST_LEAF_SPLITS,
  ST_INNER_SPLITS,
  ST_ROOT_SPLITS,
  ST_DEFERRED,
  ST_UPSERTS,
  ST_ROUTE_HITS,
  ST_ROUTE_MISSES,
  ST_NODE_READS,  // tree nodes read by lookups and transactions (the reference's traversal steps)
  ST_LEAF_STEPS,  // right-sibling steps along the leaf level (the reference's leaf traversal steps)
  ST_KEYS,  // keys added since the last bulk build (kept by stm_reset_stats); must stay last
  ST_NUM
};
constexpr int STAT_STRIPES = 256;
constexpr int STATS_WORDS = ST_NUM * STAT_STRIPES;
constexpr int STATS_RESET_WORDS = ST_KEYS * STAT_STRIPES;

STM_DEV void stat_add(unsigned long long* stats, int slot, uint32_t stripe, uint64_t v) {
  if (v) atomicAdd(stats + slot * STAT_STRIPES + (stripe & (STAT_STRIPES - 1)), (unsigned long long)v);
}

// Learned-routing directory entry r: the low fence and node id of the r-th leaf in key order
// when the directory was built, its level-1 parent then (a hint for splits) and whether the
// leaf has split since (its live high fence must then be read).  16 bytes, one load.
struct alignas(16) DirEntry {
  uint64_t fence;
  uint32_t lid;
  uint32_t pflags;  // parent id in bits 0..30 (DIR_PAR_NONE: unknown), DIR_SPLIT in bit 31
};
constexpr uint32_t DIR_SPLIT = 1u << 31;
constexpr uint32_t DIR_PAR_NONE = DIR_SPLIT - 1;

struct TreeView {
  uint64_t* keys;
  uint64_t* vals;
  NodeMeta* meta;
  TreeHeader* hdr;
  unsigned long long* stats;
  uint32_t capacity;
  uint32_t fanout;
  // Learned-routing directory hooks (null without a router): a committed leaf split marks
  // DIR_SPLIT in the leaf's directory entry so lookups re-read its live high fence.
  const uint32_t* rank_of;
  DirEntry* dir;
};

// ---------------------------------------------------------------- memory-model helpers
// The node pool is written by transactions on any XCD while other XCDs read it.  Every
// access of the transactional path to node words, lock words and the tree header is an
// agent-scope atomic (gfx950: global_load/global_store with sc1, CAS in L2), so each
// single access is coherent across the eight XCD L2 caches of the device-local pool.
// Ordering between those accesses only needs the issuing wavefront to wait for
// completion: wave_order() waits until every earlier load has returned and every earlier
// store has been acknowledged at device scope before any later access issues.  The
// agent-scope fences of the LLVM model would add an L2 write-back (buffer_wbl2 sc1) and
// invalidate (buffer_inv sc1), which only matter for non-atomic accesses.  The read-only
// lookup kernel uses plain loads and runs only when no writer is active; kernel boundaries
// order it against the transactional kernels.
// One exception inside the transactional kernels: the lane-group fast path of k_upsert_fast
// (fast_leaf_upserts in stm_tree.hip) reads and writes a leaf's fences, keys and values with
// plain loads and stores.  That is safe only because a group touches a leaf solely when it owns
// the leaf in that launch (no other merged INSERT of the batch lies between the leaf's low fence
// and its key, which the batch's sorted order proves) and the leaf's lock word still holds the
// version the lookup kernel recorded: then no other group of the launch reads or writes that
// leaf, so the plain loads return what earlier kernels left, and the launch boundary publishes
// the plain stores.  A change that lets a group of that kernel touch a leaf it does not own
// must use the agent-scope accesses and the lock word instead, as k_upsert does.
// Plain load through a pointer known to address device global memory.  Pointers read from
// memory (a batch descriptor's) are generic to the compiler, which then emits flat loads; this
// lets it use global loads with a scalar base and immediate offsets.
template <typename T>
STM_DEV T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

STM_DEV uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
STM_DEV void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
STM_DEV bool cas_u64(uint64_t* p, uint64_t expect, uint64_t desired) {
  return __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
STM_DEV void wave_order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion across this point
  __builtin_amdgcn_s_waitcnt(0);            // vmcnt(0) expcnt(0) lgkmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---------------------------------------------------------------- wavefront helpers
STM_DEV int lane_id() { return __lane_id(); }
STM_DEV uint64_t ballot(bool p) { return __ballot(p); }
STM_DEV uint32_t uniform32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
STM_DEV uint64_t uniform64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
STM_DEV uint64_t shfl64(uint64_t v, int src) {
  // ds_bpermute moves 32 bits; two of them move a u64.
  int lo = __shfl((int)(uint32_t)v, src, WAVE);
  int hi = __shfl((int)(uint32_t)(v >> 32), src, WAVE);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
STM_DEV uint64_t readlane64(uint64_t v, int src) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
STM_DEV uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__host__ __device__ inline uint32_t meta_count(uint64_t info) { return (uint32_t)info; }
__host__ __device__ inline uint32_t meta_level(uint64_t info) { return (uint32_t)(info >> 32); }
__host__ __device__ inline uint64_t make_info(uint32_t count, uint32_t level) {
  return (uint64_t)count | ((uint64_t)level << 32);
}

}  // namespace stm
