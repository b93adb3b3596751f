// stm_api.hip — the C ABI of include/stmtree.h: tree object, batch pipeline, bulk build,
// dump, structural check and statistics.
//
// Batch pipeline on the tree's stream (reference: GPU_STM_Tree::gpu_work_main,
// src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-167, and launchKernelBefore/launchKernel/
// launchKernelAfter, src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:264-332):
//   sort (key, position)  ->  per-key fold  ->  read-only lookups  ->  transactional upserts
// Batches longer than max_batch are cut into consecutive chunks, as the reference cuts
// its input file into batch_size pieces; each chunk completes before the next starts,
// so results do not depend on the chunking.  No host round trip happens inside a chunk
// (the reference copies the run count to the host between RLE and scan, :303).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/stmtree.h"
#include "stm_common.hpp"
#include "stm_internal.hpp"
#include "stm_tree.hpp"

using namespace stm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(e_ == hipErrorOutOfMemory ? STM_ERR_NOMEM : STM_ERR_DEVICE, "%s: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, count * sizeof(T));
}

// Totals of a batch stream, accumulated on the device so replays need no host round trip.
struct StreamAcc {
  unsigned long long unique;   // groups after the per-key fold
  unsigned long long bad;      // ops rejected for the reserved key
  unsigned long long batches;  // batches that ran
  unsigned long long skipped;  // batches skipped after an error
  unsigned long long sort_stalls;  // batches whose sort gave up a look-back wait
  unsigned long long fallbacks;    // batches whose normalized sort fell back to the exact passes
  unsigned long long leaf_splits;  // leaves the batches split
  unsigned long long pad[1];
};

constexpr uint32_t DESC_TABLE = 4096;  // descriptors uploaded per window of a stream
constexpr int GRAPH_CLASSES = 12;      // size classes 2^10, 2^12, ..., up to max_batch

// First node of every replayed graph: takes the next descriptor of the window.  Once a batch
// has failed (reserved key, full node pool) the later batches of the stream run empty, so the
// tree holds exactly the batches up to the failing one.
__global__ void k_advance(const BatchDesc* __restrict__ table, uint32_t* __restrict__ cursor,
                          BatchDesc* __restrict__ live, const StreamAcc* __restrict__ acc,
                          const TreeHeader* __restrict__ hdr) {
  if (threadIdx.x != 0) return;
  const uint32_t c = cursor[0];
  BatchDesc d = table[c];
  if (acc->bad || acc->sort_stalls || hdr->error) d.n = 0;
  *live = d;
  cursor[0] = c + 1;
}

// Last node of every replayed graph.
__global__ void k_epilogue(const BatchCounters* __restrict__ cnt, const BatchDesc* __restrict__ live,
                           StreamAcc* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  if (live->n == 0) {
    acc->skipped += 1;
    return;
  }
  acc->unique += cnt->n_unique;
  acc->bad += cnt->n_bad;
  acc->sort_stalls += cnt->sort_stall;
  acc->fallbacks += cnt->sort_fallback;
  acc->leaf_splits += cnt->leaf_splits;
  acc->batches += 1;
}

// First node of a back half: the front half of this batch may have run beside the back half of
// a batch that failed, so the failure is checked again here, after every earlier back half
// (same stream).  A batch after a failure keeps no groups: its lookups and upserts do nothing.
__global__ void k_gate(const StreamAcc* __restrict__ acc, const TreeHeader* __restrict__ hdr,
                       BatchDesc* __restrict__ live, BatchCounters* __restrict__ cnt) {
  if (threadIdx.x != 0) return;
  if (live->n == 0) return;
  if (acc->bad || acc->sort_stalls || hdr->error) {
    live->n = 0;
    cnt->n_unique = 0;
    cnt->n_upserts = 0;
    cnt->n_fanout = 0;
    cnt->n_deferred = 0;
  }
}

// Plain streaming copy of n16 16-byte words: the known-size reference kernel for HBM traffic
// counters and the measured bandwidth ceiling.  One word per thread over a grid that covers the
// whole copy: on gfx950 this outruns both a resident grid-stride loop and several words in
// flight per thread (6.2 TB/s read + write against 4.7-5.6 TB/s at 4 GiB).  The stride loop
// only serves copies beyond COPY_MAX_BLOCKS blocks.
constexpr uint32_t COPY_MAX_BLOCKS = 1u << 24;
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                uint64_t n16) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256ull)
    dst[i] = src[i];
}

struct Scratch {
  uint64_t n = 0;
  uint64_t *in_keys = nullptr, *in_vals = nullptr, *out_vals = nullptr;
  uint8_t *in_ops = nullptr, *out_found = nullptr;
  uint64_t *skeys_tmp = nullptr, *skeys = nullptr;
  uint32_t *sidx_tmp = nullptr, *sidx = nullptr;
  uint32_t* hist = nullptr;
  void* fold_tiles = nullptr;
  uint64_t *ukey = nullptr, *uval = nullptr;
  uint8_t* uop = nullptr;
  uint32_t *gstart = nullptr, *npend = nullptr, *dlist = nullptr;
  uint32_t* slist = nullptr;  // merged INSERTs the fast path hands back
  uint32_t* scount = nullptr;  // per fast-path wavefront: how many it handed back
  uint32_t* ilist = nullptr;  // merged INSERTs of the batch
  InsHint* ihint = nullptr;   // the lookup kernel's record of every merged INSERT
  uint64_t* fanout = nullptr; // answer chunks of hot GET groups (fanout_capacity(n))
  BatchCounters* counters = nullptr;
  BatchCounters* h_counters = nullptr;  // pinned host copy for the per-batch readback
  // range scans
  uint64_t *soff = nullptr, *scan_tiles = nullptr;
  int32_t* lastins = nullptr;
  // batch streams replayed from captured graphs
  BatchDesc* desc_table = nullptr;  // DESC_TABLE entries
  BatchDesc* desc_live = nullptr;   // the batch the current replay works on
  uint32_t* cursor = nullptr;       // next desc_table entry
  StreamAcc* acc = nullptr;

  // Every device array of the scratch for batches of up to cap ops, in one fixed order:
  // take(&pointer, bytes) provides each one (allocation, or a place in a fixed region).
  template <typename Take>
  hipError_t layout(uint64_t cap, bool staging, Take&& take) {
    n = cap;
    hipError_t e;
#define A_(p, c) \
    if ((e = take((void**)&p, sizeof(*p) * std::max<uint64_t>((uint64_t)(c), 1))) != hipSuccess) return e
    if (staging) {
      A_(in_keys, cap); A_(in_vals, cap); A_(in_ops, cap); A_(out_vals, cap); A_(out_found, cap);
    }
    A_(skeys_tmp, cap); A_(skeys, cap); A_(sidx_tmp, cap); A_(sidx, cap);
    A_(hist, sort_hist_words(cap));
    uint8_t* ft = nullptr;
    A_(ft, fold_scratch_bytes(cap));
    fold_tiles = ft;
    A_(ukey, cap); A_(uval, cap); A_(uop, cap); A_(gstart, cap); A_(npend, cap);
    A_(dlist, cap);
    A_(ihint, cap);
    A_(slist, cap);
    A_(scount, FAST_WAVES_MAX);
    A_(ilist, cap);
    A_(fanout, fanout_capacity(cap));
    A_(counters, 1);
    A_(desc_live, 1);
    if (staging) {
      A_(desc_table, DESC_TABLE); A_(cursor, 1); A_(acc, 1);
      A_(soff, cap); A_(lastins, cap); A_(scan_tiles, scan_tile_words(cap));
    }
#undef A_
    return hipSuccess;
  }

  hipError_t alloc(uint64_t cap, bool staging) {
    hipError_t e = layout(cap, staging, [](void** p, size_t bytes) { return hipMalloc(p, bytes); });
    if (e != hipSuccess) return e;
    if ((e = sort_scratch_init(hist)) != hipSuccess) return e;
    return hipHostMalloc((void**)&h_counters, sizeof(BatchCounters), hipHostMallocCoherent);
  }

  // Lays the arrays out back to back (256-byte aligned) from base; with base null only
  // counts.  Returns the bytes the layout spans.
  size_t carve(uint64_t cap, uint8_t* base) {
    size_t at = 0;
    (void)layout(cap, false, [&](void** p, size_t bytes) {
      at = (at + 255) / 256 * 256;
      *p = base ? base + at : nullptr;
      at += bytes;
      return hipSuccess;
    });
    return at;
  }
  void release() {
    void* ps[] = {in_keys, in_vals, out_vals, in_ops, out_found, skeys_tmp, skeys, sidx_tmp, sidx,
                  hist, fold_tiles, ukey, uval, uop, gstart, npend, dlist, ihint, slist, scount, ilist, fanout, counters,
                  desc_table, desc_live, cursor, acc, soff, lastins, scan_tiles};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    if (h_counters) (void)hipHostFree(h_counters);
    *this = Scratch();
  }
};

}  // namespace

struct stm_tree {
  stm_config cfg{};
  hipStream_t stream = nullptr;
  TreeView view{};
  uint64_t capacity = 0;
  Scratch scr;
  stm_stats st{};
  hipEvent_t ev_a[STM_NUM_PHASES] = {};
  hipEvent_t ev_b[STM_NUM_PHASES] = {};
  hipEvent_t ev_begin[STM_NUM_PHASES] = {};  // the event a phase's time starts from
  bool phase_used[STM_NUM_PHASES] = {};
  // learned routing
  RouteState route;
  bool route_alloc = false;
  bool route_valid = false;       // directory + model describe a recent tree
  bool route_warm = false;        // the segment models were trained on this tree
  uint32_t batches_since_train = 0;
  uint64_t splits_since_train = 0;  // leaf splits counted since the last directory rebuild
  uint64_t dir_leaves = 0;          // leaves of the directory at the last rebuild
  // Batch streams (stm_run_batches_device) run each batch as two captured halves on two
  // scratch sets, so the front half (descriptor, sort, fold) of batch i + 1 runs on
  // stream_front while the back half (lookups, upserts, counters) of batch i runs on `stream`.
  // One graph per half, size class and scratch set (null until first needed).
  Scratch scr2;
  bool scr2_alloc = false;
  hipStream_t stream_front = nullptr;
  hipEvent_t ev_front[2] = {};  // front half of the slot's latest batch done
  hipEvent_t ev_back[2] = {};   // back half of the slot's latest batch done (its scratch is free)
  hipEvent_t ev_sync = nullptr;
  hipGraphExec_t graphs_front[GRAPH_CLASSES][2] = {};
  hipGraphExec_t graphs_back[GRAPH_CLASSES][2] = {};
  // Batches above 2^18 ops sort normalized keys unless exact_left > 0: a batch whose normalized
  // sort had to fall back (long buckets of distinct keys: hot keys with neighbours) makes the
  // next EXACT_HOLD batches use the eight exact passes directly.  graphs_exact: the form the
  // captured graphs hold.
  uint32_t exact_left = 0;
  bool graphs_exact = false;
  uint64_t sort_fallbacks = 0;  // batches whose normalized sort fell back (stm_stats)
};

namespace {

constexpr uint32_t EXACT_HOLD = 16;

uint64_t capacity_for(uint64_t max_keys, uint32_t F) {
  const uint64_t half = F / 2;
  uint64_t leaves = (max_keys + half - 1) / half + 1;
  uint64_t cap = leaves + leaves / (half - 1) + 64;
  cap += cap / 10 + 1024;
  return cap;
}

int set_device(stm_tree* t) {
  HIP_TRY(hipSetDevice(t->cfg.device));
  return STM_OK;
}

void phase_begin(stm_tree* t, int p) {
  if (!t->cfg.timing) return;
  (void)hipEventRecord(t->ev_a[p], t->stream);
  t->ev_begin[p] = t->ev_a[p];
}
void phase_end(stm_tree* t, int p) {
  if (!t->cfg.timing) return;
  (void)hipEventRecord(t->ev_b[p], t->stream);
  t->phase_used[p] = true;
}
// Phase `from` ends where phase `to` begins: one event marks both.  Each recorded event sits
// between two kernels and holds the second back by a few microseconds, so a boundary costs one
// event instead of two.
void phase_switch(stm_tree* t, int from, int to) {
  if (!t->cfg.timing) return;
  phase_end(t, from);
  t->ev_begin[to] = t->ev_b[from];
}
void phase_collect(stm_tree* t) {
  if (!t->cfg.timing) return;
  for (int p = 0; p < STM_NUM_PHASES; ++p) {
    if (!t->phase_used[p]) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t->ev_begin[p], t->ev_b[p]) == hipSuccess) {
      t->st.phase_ms[p] += ms;
      t->st.phase_launches[p] += 1;
    }
    t->phase_used[p] = false;
  }
}

// Sum of the stripes of every device counter (the stream is synchronised first).
int read_counters(stm_tree* t, uint64_t out[ST_NUM]) {
  std::vector<unsigned long long> c(STATS_WORDS);
  HIP_TRY(hipStreamSynchronize(t->stream));
  HIP_TRY(hipMemcpy(c.data(), t->view.stats, sizeof(unsigned long long) * STATS_WORDS,
                    hipMemcpyDeviceToHost));
  for (int k = 0; k < ST_NUM; ++k) {
    uint64_t v = 0;
    for (int j = 0; j < STAT_STRIPES; ++j) v += c[(size_t)k * STAT_STRIPES + j];
    out[k] = v;
  }
  return STM_OK;
}

int read_header(stm_tree* t, TreeHeader* h) {
  HIP_TRY(hipMemcpyAsync(h, t->view.hdr, sizeof(TreeHeader), hipMemcpyDeviceToHost, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  return STM_OK;
}

int error_bits_status(stm_tree* t, uint64_t error) {
  if (error & ERR_POOL_FULL)
    return fail(STM_ERR_NOMEM, "node pool exhausted (capacity %llu nodes); raise max_keys",
                (unsigned long long)t->capacity);
  if (error & ERR_LIVELOCK) return fail(STM_ERR_DEVICE, "serial upsert pass could not commit");
  return STM_OK;
}

int check_device_error(stm_tree* t) {
  TreeHeader h;
  int rc = read_header(t, &h);
  if (rc) return rc;
  return error_bits_status(t, h.error);
}

// Writes an empty tree: one empty leaf as root.
int init_empty(stm_tree* t) {
  NodeMeta m{};
  m.lock = 0;
  m.info = 0;
  m.low = 0;
  m.high = KEY_EMPTY;
  m.next = NODE_NONE;
  HIP_TRY(hipMemsetAsync(t->view.keys, 0xff, sizeof(uint64_t) * t->cfg.fanout, t->stream));
  HIP_TRY(hipMemsetAsync(t->view.vals, 0, sizeof(uint64_t) * t->cfg.fanout, t->stream));
  HIP_TRY(hipMemcpyAsync(t->view.meta, &m, sizeof(m), hipMemcpyHostToDevice, t->stream));
  TreeHeader h{};
  h.lock = 0;
  h.root = 0;
  h.height = 1;
  h.n_nodes = 1;
  h.capacity = t->capacity;
  h.error = 0;
  h.n_keys = 0;
  HIP_TRY(hipMemcpyAsync(t->view.hdr, &h, sizeof(h), hipMemcpyHostToDevice, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  return STM_OK;
}

void drop_graphs(stm_tree* t) {
  for (int c = 0; c < GRAPH_CLASSES; ++c)
    for (int k = 0; k < 2; ++k) {
      hipGraphExec_t* gs[] = {&t->graphs_front[c][k], &t->graphs_back[c][k]};
      for (hipGraphExec_t* g : gs) {
        if (*g) (void)hipGraphExecDestroy(*g);
        *g = nullptr;
      }
    }
}

int alloc_router(stm_tree* t) {
  if (t->route_alloc) return STM_OK;
  RouteState& r = t->route;
  const uint64_t cap = t->capacity;
  if (dalloc(&r.hdr, 1) || dalloc(&r.fences, cap + 1) || dalloc(&r.lids, cap) ||
      dalloc(&r.segs, ROUTE_SEGMENTS) || dalloc(&r.dir, cap + 1) || dalloc(&r.rank_of, cap) ||
      dalloc(&r.lows, cap) || dalloc(&r.lows_tmp, cap) || dalloc(&r.sorted_lows, cap) ||
      dalloc(&r.order, cap) || dalloc(&r.order_tmp, cap) || dalloc(&r.hist, sort_hist_words(cap)) ||
      dalloc(&r.dir_alt, cap + 1))
    return fail(STM_ERR_NOMEM, "learned-routing buffers for %llu nodes", (unsigned long long)cap);
  if (sort_scratch_init(r.hist) != hipSuccess) return fail(STM_ERR_DEVICE, "learned-routing sort scratch");
  t->route_alloc = true;
  t->view.rank_of = r.rank_of;
  t->view.dir = r.dir;
  drop_graphs(t);  // captured with the previous TreeView
  return STM_OK;
}

void free_router(stm_tree* t) {
  RouteState& r = t->route;
  void* ps[] = {r.hdr, r.fences, r.lids, r.segs, r.dir, r.rank_of, r.lows, r.lows_tmp,
                r.sorted_lows, r.order, r.order_tmp, r.hist, r.dir_alt};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  t->route = RouteState();
  t->view.rank_of = nullptr;
  t->view.dir = nullptr;
  t->route_alloc = false;
  t->route_valid = false;
  t->route_warm = false;
}

// Rebuilds the leaf directory from the tree and trains the router (phase "train").  A
// periodic retrain of models already trained on this tree starts from their parameters and
// runs an eighth of the steps (at least 16): a few batches move few leaves.
int train_router(stm_tree* t, uint32_t steps, bool periodic) {
  int rc = alloc_router(t);
  if (rc) return rc;
  TreeHeader h;
  rc = read_header(t, &h);
  if (rc) return rc;
  const uint32_t nn = (uint32_t)std::min<uint64_t>(h.n_nodes, t->capacity);
  const bool index = t->cfg.route == STM_ROUTE_INDEX;
  const bool warm = periodic && t->route_warm;
  // A periodic retrain fine-tunes the converged segment models over the rewritten directory:
  // a few Adam steps (9 of the default 300).  At 2^26 Gamma keys and 2^22-op batches, 37 steps
  // kept 91% of lookups in the predicted window and 9 keep 84%, with the same lookup time (a
  // sorted span reuses its leaf, and a missed window costs one more window read), while the
  // training time per retrain drops by three quarters.
  if (warm) steps = std::max<uint32_t>(steps / 32, 8);
  if (index) steps = 0;  // the static index only needs the directory
  // Adam's first steps move every parameter by about the learning rate whatever the
  // gradient's size, so fine-tuning a converged model uses a tenth of it.
  const float lr = warm ? 0.002f : 0.02f;
  phase_begin(t, STM_PHASE_TRAIN);
  HIP_TRY(route_build(t->view, t->route, nn, steps, lr, steps > 0, warm, t->stream));
  t->route_warm = steps > 0 || index;
  phase_end(t, STM_PHASE_TRAIN);
  if (t->cfg.timing) {  // every retrain is counted, also between graph replays
    HIP_TRY(hipStreamSynchronize(t->stream));
    phase_collect(t);
  }
  t->route_valid = true;
  t->batches_since_train = 0;
  t->splits_since_train = 0;
  RouteHeader rh{};
  HIP_TRY(hipMemcpyAsync(&rh, t->route.hdr, sizeof(rh), hipMemcpyDeviceToHost, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  t->dir_leaves = rh.L;
  return STM_OK;
}

// The launch sequence of one batch: sort -> fold -> lookups -> upserts.  With in.desc the
// batch (its arrays and its length, at most in.n) is read on the device, which is what a
// captured graph replays; otherwise in.n and the pointers describe it.  No host round trip.
struct PipeIn {
  const uint64_t* keys;
  const uint64_t* vals;
  const uint8_t* ops;
  uint32_t n;
  uint64_t* out_vals;
  uint8_t* out_found;
  const BatchDesc* desc;
  // scan mode: op 3 is SCAN(key, L = vals[i]) with output into scan_keys / scan_vals
  bool scans;
  uint64_t* scan_keys;
  uint64_t* scan_vals;
  uint64_t scan_cap;
  // optional: the last upsert kernel copies the batch counters here (pinned host memory)
  BatchCounters* host_counters;
};

// Front half of a batch: sort and fold.  Reads only the batch and writes only the scratch and
// the outputs of ops the fold answers, never the tree, so it may run on `st` beside the back
// half of the previous batch (on another scratch).
hipError_t enqueue_front(stm_tree* t, Scratch& s, const PipeIn& in, bool timed, bool exact,
                         hipStream_t st) {
  hipError_t e;
  if (timed) phase_begin(t, STM_PHASE_SORT);
  e = radix_sort_pairs_desc(in.keys, in.ops, in.desc, in.n, s.skeys_tmp, s.skeys, s.sidx_tmp, s.sidx, 0,
                            64, s.hist, (int)t->cfg.sort_tile_items, exact, st);
  if (e != hipSuccess) return e;
  if (timed) phase_switch(t, STM_PHASE_SORT, STM_PHASE_MERGE);

  FoldArgs fa{};
  fa.skeys = s.skeys;
  fa.sidx = s.sidx;
  fa.vals = in.vals;
  fa.ops = in.ops;
  fa.n = in.n;
  fa.desc = in.desc;
  fa.out_vals = in.out_vals;
  fa.out_found = in.out_found;
  fa.ukey = s.ukey;
  fa.uval = s.uval;
  fa.uop = s.uop;
  fa.gstart = s.gstart;
  fa.npend = s.npend;
  fa.counters = s.counters;
  fa.tile_scratch = s.fold_tiles;
  fa.lastins = in.scans ? s.lastins : nullptr;
  fa.ilist = s.ilist;
  fa.sort_stall = sort_stall_word(s.hist);
  fa.sort_exact = sort_exact_word(s.hist);
  if ((e = fold_batch(fa, st)) != hipSuccess) return e;
  if (timed) phase_switch(t, STM_PHASE_MERGE, STM_PHASE_LOOKUP);
  return hipSuccess;
}

// Back half: lookups, scans, upserts (and the counters' copy to the host when asked), on the
// tree's stream.
hipError_t enqueue_back(stm_tree* t, Scratch& s, const PipeIn& in, bool timed) {
  hipStream_t st = t->stream;
  hipError_t e;
  ScanArgs sa{};
  if (in.scans) {
    sa.keys = in.keys;
    sa.vals = in.vals;
    sa.ops = in.ops;
    sa.n = in.n;
    sa.desc = in.desc;
    sa.out_vals = in.out_vals;
    sa.out_found = in.out_found;
    sa.scan_keys = in.scan_keys;
    sa.scan_vals = in.scan_vals;
    sa.scan_cap = in.scan_cap;
    sa.ukey = s.ukey;
    sa.uop = s.uop;
    sa.gstart = s.gstart;
    sa.npend = s.npend;
    sa.sidx = s.sidx;
    sa.lastins = s.lastins;
    sa.counters = s.counters;
    sa.soff = s.soff;
    sa.tile_sums = s.scan_tiles;
  }

  LookupArgs la{};
  if (t->cfg.route != STM_ROUTE_DESCENT) {
    la.route_hdr = t->route.hdr;
    la.fences = t->route.fences;
    la.lids = t->route.lids;
    la.segs = t->route.segs;
    la.index = t->cfg.route == STM_ROUTE_INDEX ? 1u : 0u;
  }
  la.ukey = s.ukey;
  la.uop = s.uop;
  la.gstart = s.gstart;
  la.npend = s.npend;
  la.sidx = s.sidx;
  la.counters = s.counters;
  la.n_max = in.n;
  la.chunk = t->cfg.ops_per_group;
  la.lanes = t->cfg.lanes_per_op;
  la.out_vals = in.out_vals;
  la.out_found = in.out_found;
  la.desc = in.desc;
  la.ihint = s.ihint;
  la.gval = s.uval;  // GET groups carry no merged value: a hot group's value goes there
  la.fanout = s.fanout;
  la.fanout_cap = fanout_capacity(in.n);
  if ((e = launch_lookup(t->view, la, st)) != hipSuccess) return e;

  if (in.scans) {  // on the tree as it was before the batch: before the upsert kernel
    if (timed) phase_switch(t, STM_PHASE_LOOKUP, STM_PHASE_SCAN);
    if ((e = launch_scan_offsets(sa, st)) != hipSuccess) return e;
    if ((e = launch_range_scan(t->view, sa, st)) != hipSuccess) return e;
    if (timed) phase_switch(t, STM_PHASE_SCAN, STM_PHASE_UPSERT);
  } else if (timed) {
    phase_switch(t, STM_PHASE_LOOKUP, STM_PHASE_UPSERT);
  }

  UpsertArgs ua{};
  ua.ukey = s.ukey;
  ua.uval = s.uval;
  ua.uop = s.uop;
  ua.gstart = s.gstart;
  ua.npend = s.npend;
  ua.sidx = s.sidx;
  ua.counters = s.counters;
  ua.dlist = s.dlist;
  ua.max_attempts = t->cfg.max_attempts;
  // Each wavefront owns the groups of upsert_ops_per_wave batch slots; at most 8192 wavefronts
  // (every SIMD of the chip fully occupied).
  const uint64_t waves = ((uint64_t)in.n + t->cfg.upsert_ops_per_wave - 1) / t->cfg.upsert_ops_per_wave;
  ua.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((waves + 3) / 4, 2048));
  ua.out_vals = in.out_vals;
  ua.out_found = in.out_found;
  ua.desc = in.desc;
  ua.ihint = s.ihint;
  ua.ilist = s.ilist;
  ua.slist = s.slist;
  ua.scount = s.scount;
  ua.lanes = t->cfg.upsert_lanes;
  ua.fanout = s.fanout;
  ua.fanout_cap = fanout_capacity(in.n);
  ua.gval = s.uval;
  ua.host_counters = in.host_counters;
  if ((e = launch_upsert(t->view, ua, st)) != hipSuccess) return e;
  if (timed) phase_end(t, STM_PHASE_UPSERT);
  if (in.scans && (e = launch_scan_answers(sa, st)) != hipSuccess) return e;
  return hipSuccess;
}

hipError_t enqueue_pipeline(stm_tree* t, Scratch& s, const PipeIn& in, bool timed, bool exact) {
  hipError_t e = enqueue_front(t, s, in, timed, exact, t->stream);
  if (e != hipSuccess) return e;
  return enqueue_back(t, s, in, timed);
}

// Learned routing and the static index: (re)build the directory, and retrain, when it is
// missing or stale.
// Also after split-heavy batches: once the leaves split since the last rebuild reach
// retrain_splits (a 32nd of the directory's leaves by default), lookups would walk right
// siblings from too many stale entries.
int maybe_retrain(stm_tree* t) {
  if (t->cfg.route == STM_ROUTE_DESCENT) return STM_OK;
  const uint64_t split_limit =
      t->cfg.retrain_splits ? t->cfg.retrain_splits : std::max<uint64_t>(1, t->dir_leaves / 32);
  if (t->route_valid && t->batches_since_train < t->cfg.retrain_every &&
      t->splits_since_train < split_limit)
    return STM_OK;
  return train_router(t, t->cfg.train_steps, true);
}

// Scan output area of a chunk (device pointers); null = op 3 runs as GET.
struct ScanOut {
  uint64_t* keys;
  uint64_t* vals;
  uint64_t cap;
  uint64_t used;  // set by run_chunk: the sum of the chunk's scan lengths
};

// One chunk of at most max_batch ops, all pointers on the device; waits for its completion.
// The batch scratch is the tree's own unless `scr` points at another layout of it.
int run_chunk(stm_tree* t, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
              uint32_t n, uint64_t* out_vals, uint8_t* out_found, ScanOut* so = nullptr,
              Scratch* scr = nullptr) {
  Scratch& s = scr ? *scr : t->scr;
  hipStream_t st = t->stream;
  int rc = maybe_retrain(t);
  if (rc) return rc;
  PipeIn in{keys, vals, ops, n, out_vals, out_found, nullptr, false, nullptr, nullptr, 0, nullptr};
  if (so) {
    in.scans = true;
    in.scan_keys = so->keys;
    in.scan_vals = so->vals;
    in.scan_cap = so->cap;
  } else {
    in.host_counters = s.h_counters;  // the scan answers kernel runs after the upsert kernels
  }
  const bool exact = t->exact_left > 0;
  HIP_TRY(enqueue_pipeline(t, s, in, true, exact));
  // One readback per batch, into pinned memory: counters and the tree's error bits.
  if (so)
    HIP_TRY(hipMemcpyAsync(s.h_counters, s.counters, sizeof(BatchCounters), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const BatchCounters bc = *s.h_counters;
  phase_collect(t);
  if (bc.sort_stall) return fail(STM_ERR_DEVICE, "batch sort: a look-back wait gave up");
  if (bc.sort_fallback) {
    t->exact_left = EXACT_HOLD;
    t->sort_fallbacks += 1;
  } else if (exact) {
    t->exact_left -= 1;
  }
  t->st.batches += 1;
  t->st.ops += n;
  t->st.unique_ops += bc.n_unique;
  t->batches_since_train += 1;
  t->splits_since_train += bc.leaf_splits;
  if (bc.n_bad) return fail(STM_ERR_INVALID, "batch holds the reserved key 0xffffffffffffffff");
  if (so) {
    so->used = bc.scan_total;
    if (bc.scan_bad)
      return fail(STM_ERR_INVALID, "%u scan ops have a length outside [1, %u] or overflow the "
                  "scan output area (%llu pairs)", bc.scan_bad, SCAN_MAX_LEN,
                  (unsigned long long)so->cap);
  }
  return error_bits_status(t, bc.hdr_error);
}

// ------------------------------------------------------------------ graph replay of streams
uint32_t class_nmax(const stm_tree* t, int c) {
  const uint64_t v = 1024ull << (2 * c);
  return (uint32_t)std::min<uint64_t>(v, t->cfg.max_batch);
}

int class_of(const stm_tree* t, uint32_t n) {
  for (int c = 0; c < GRAPH_CLASSES; ++c)
    if (class_nmax(t, c) >= n) return c;
  return GRAPH_CLASSES - 1;  // max_batch <= 2^30 fits class 10
}

// The second scratch set, the front stream and the pipeline's events, made on first use.
int ensure_pipeline(stm_tree* t) {
  if (t->scr2_alloc) return STM_OK;
  hipError_t e = t->scr2.alloc(t->cfg.max_batch, false);
  if (e == hipSuccess) e = hipStreamCreate(&t->stream_front);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipEventCreateWithFlags(&t->ev_front[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ev_back[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ev_sync, hipEventDisableTiming);
  if (e != hipSuccess) return fail(STM_ERR_NOMEM, "stream pipeline set-up: %s", hipGetErrorString(e));
  t->scr2_alloc = true;
  return STM_OK;
}

Scratch& slot_scratch(stm_tree* t, int k) { return k ? t->scr2 : t->scr; }

// Captures the two halves of size class c on scratch set k:
//   front: advance (next descriptor of the window into the set's live descriptor) -> sort -> fold
//   back:  gate -> lookups -> upserts -> epilogue (counters into the stream's accumulator)
// The descriptor window, its cursor and the accumulator are the first set's.
int capture_pair(stm_tree* t, int c, int k) {
  if (t->graphs_front[c][k] && t->graphs_back[c][k]) return STM_OK;
  Scratch& s0 = t->scr;
  Scratch& s = slot_scratch(t, k);
  PipeIn in{nullptr, nullptr, nullptr, class_nmax(t, c), nullptr, nullptr, s.desc_live,
            false, nullptr, nullptr, 0, nullptr};
  for (int half = 0; half < 2; ++half) {
    hipStream_t st = half ? t->stream : t->stream_front;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e;
    if (half == 0) {
      k_advance<<<1, 64, 0, st>>>(s0.desc_table, s0.cursor, s.desc_live, s0.acc, t->view.hdr);
      e = hipGetLastError();
      if (e == hipSuccess) e = enqueue_front(t, s, in, false, t->graphs_exact, st);
    } else {
      k_gate<<<1, 64, 0, st>>>(s0.acc, t->view.hdr, s.desc_live, s.counters);
      e = hipGetLastError();
      if (e == hipSuccess) e = enqueue_back(t, s, in, false);
      if (e == hipSuccess) {
        k_epilogue<<<1, 64, 0, st>>>(s.counters, s.desc_live, s0.acc);
        e = hipGetLastError();
      }
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t& x = half ? t->graphs_back[c][k] : t->graphs_front[c][k];
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      x = nullptr;
      return fail(STM_ERR_DEVICE, "graph capture of size class %d: %s", c, hipGetErrorString(e));
    }
  }
  return STM_OK;
}

int capture_class(stm_tree* t, int c) {
  int rc = ensure_pipeline(t);
  for (int k = 0; k < 2 && rc == STM_OK; ++k) rc = capture_pair(t, c, k);
  return rc;
}

int run_stream(stm_tree* t, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
               const uint64_t* offsets, size_t nb, uint64_t* out_vals, uint8_t* out_found) {
  Scratch& s = t->scr;
  hipStream_t st = t->stream;
  const uint64_t mb = t->cfg.max_batch;
  // Batches cut into chunks of at most max_batch ops, in order.
  std::vector<BatchDesc> descs;
  std::vector<uint8_t> cls;
  for (size_t b = 0; b < nb; ++b) {
    for (uint64_t off = offsets[b]; off < offsets[b + 1]; off += mb) {
      const uint32_t m = (uint32_t)std::min<uint64_t>(mb, offsets[b + 1] - off);
      descs.push_back(BatchDesc{keys + off, vals + off, ops + off, out_vals + off, out_found + off, m,
                                0, nullptr, nullptr, 0});
      cls.push_back((uint8_t)class_of(t, m));
    }
  }
  if (descs.empty()) return STM_OK;
  int rc = maybe_retrain(t);  // learned routing: the router's buffers exist before capture
  if (rc) return rc;
  if (t->graphs_exact != (t->exact_left > 0)) {  // graphs of the other sort form
    drop_graphs(t);
    t->graphs_exact = t->exact_left > 0;
  }
  for (uint8_t c : cls) {
    rc = capture_class(t, c);
    if (rc) return rc;
  }
  // Batch i runs on scratch set i % 2: its front half on stream_front once the back half of
  // batch i - 2 (the set's previous user) is done, its back half on the tree's stream once its
  // front half is done.  Back halves run in batch order on one stream, so each sees the tree
  // after the previous batch; the learned router retrains on that stream between them.
  hipStream_t sf = t->stream_front;
  bool used[2] = {false, false};
  HIP_TRY(hipMemsetAsync(s.acc, 0, sizeof(StreamAcc), st));
  for (size_t w0 = 0; w0 < descs.size(); w0 += DESC_TABLE) {
    const size_t w1 = std::min<size_t>(descs.size(), w0 + DESC_TABLE);
    // The window's table and cursor are replaced once every front half of the previous window
    // has taken its descriptor.
    HIP_TRY(hipEventRecord(t->ev_sync, sf));
    HIP_TRY(hipStreamWaitEvent(st, t->ev_sync, 0));
    HIP_TRY(hipMemcpyAsync(s.desc_table, descs.data() + w0, (w1 - w0) * sizeof(BatchDesc),
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(s.cursor, 0, sizeof(uint32_t), st));
    HIP_TRY(hipEventRecord(t->ev_sync, st));
    HIP_TRY(hipStreamWaitEvent(sf, t->ev_sync, 0));
    for (size_t i = w0; i < w1; ++i) {
      const int k = (int)(i & 1);
      if (used[k]) HIP_TRY(hipStreamWaitEvent(sf, t->ev_back[k], 0));
      HIP_TRY(hipGraphLaunch(t->graphs_front[cls[i]][k], sf));
      HIP_TRY(hipEventRecord(t->ev_front[k], sf));
      rc = maybe_retrain(t);
      if (rc) return rc;
      HIP_TRY(hipStreamWaitEvent(st, t->ev_front[k], 0));
      HIP_TRY(hipGraphLaunch(t->graphs_back[cls[i]][k], st));
      HIP_TRY(hipEventRecord(t->ev_back[k], st));
      used[k] = true;
      t->st.batches += 1;
      t->st.ops += descs[i].n;
      t->batches_since_train += 1;
    }
  }
  StreamAcc acc;
  HIP_TRY(hipEventRecord(t->ev_sync, sf));
  HIP_TRY(hipStreamWaitEvent(st, t->ev_sync, 0));
  HIP_TRY(hipMemcpyAsync(&acc, s.acc, sizeof(acc), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  t->st.unique_ops += acc.unique;
  t->sort_fallbacks += acc.fallbacks;
  t->splits_since_train += acc.leaf_splits;
  if (acc.fallbacks) t->exact_left = EXACT_HOLD;
  else if (t->graphs_exact) t->exact_left -= (uint32_t)std::min<uint64_t>(t->exact_left, acc.batches);
  if (acc.sort_stalls)
    return fail(STM_ERR_DEVICE, "a batch sort gave up a look-back wait (%llu later batches skipped)",
                acc.skipped);
  if (acc.bad)
    return fail(STM_ERR_INVALID, "a batch holds the reserved key 0xffffffffffffffff (%llu later batches skipped)",
                acc.skipped);
  rc = check_device_error(t);
  if (rc && acc.skipped) {
    char buf[64];
    snprintf(buf, sizeof(buf), " (%llu later batches skipped)", acc.skipped);
    g_last_error += buf;
  }
  return rc;
}

int validate_host_batch(const uint64_t* keys, const uint8_t* ops, size_t n, bool need_vals,
                        const uint64_t* vals) {
  bool any_insert = false;
  for (size_t i = 0; i < n; ++i) {
    if (keys[i] == STM_KEY_RESERVED)
      return fail(STM_ERR_INVALID, "op %zu uses the reserved key 0xffffffffffffffff", i);
    if (ops) {
      if (ops[i] > STM_OP_RANGE) return fail(STM_ERR_INVALID, "op %zu has op code %u", i, ops[i]);
      any_insert |= ops[i] == STM_OP_INSERT;
    }
  }
  if (need_vals && (any_insert || !ops) && !vals)
    return fail(STM_ERR_INVALID, "inserts need a value array");
  return STM_OK;
}

int run_device(stm_tree* t, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
               size_t n, uint64_t* out_vals, uint8_t* out_found) {
  const uint64_t mb = t->cfg.max_batch;
  for (size_t off = 0; off < n; off += mb) {
    const uint32_t m = (uint32_t)std::min<uint64_t>(mb, n - off);
    int rc = run_chunk(t, keys + off, vals ? vals + off : nullptr, ops + off, m, out_vals + off,
                       out_found + off);
    if (rc) return rc;
  }
  return STM_OK;
}

// Bottom-up build of an empty tree from n device-resident pairs.
int build_from_device(stm_tree* t, const uint64_t* keys, const uint64_t* vals, size_t n) {
  hipStream_t st = t->stream;
  const uint32_t fill = t->cfg.leaf_fill;
  if (n > 0x7fffffffull) return fail(STM_ERR_INVALID, "bulk load of %zu keys exceeds 2^31-1", n);
  Scratch tmp;
  hipError_t e = tmp.alloc(std::max<uint64_t>(n, 1), false);
  if (e != hipSuccess) {
    tmp.release();
    return fail(STM_ERR_NOMEM, "bulk-load scratch for %zu keys: %s", n, hipGetErrorString(e));
  }
  int rc = STM_OK;
  do {
    uint32_t U = 0;
    if (n > 0) {
      hipError_t r1 = radix_sort_pairs(keys, tmp.skeys_tmp, tmp.skeys, tmp.sidx_tmp, tmp.sidx,
                                       (uint32_t)n, 0, 64, tmp.hist, st);
      if (r1 != hipSuccess) { rc = fail(STM_ERR_DEVICE, "sort: %s", hipGetErrorString(r1)); break; }
      FoldArgs fa{};
      fa.skeys = tmp.skeys;
      fa.sidx = tmp.sidx;
      fa.vals = vals;
      fa.ops = nullptr;
      fa.n = (uint32_t)n;
      fa.ukey = tmp.ukey;
      fa.uval = tmp.uval;
      fa.uop = tmp.uop;
      fa.counters = tmp.counters;
      fa.tile_scratch = tmp.fold_tiles;
      fa.sort_stall = sort_stall_word(tmp.hist);
      r1 = fold_batch(fa, st);
      if (r1 != hipSuccess) { rc = fail(STM_ERR_DEVICE, "fold: %s", hipGetErrorString(r1)); break; }
      BatchCounters bc;
      if (hipMemcpyAsync(&bc, tmp.counters, sizeof(bc), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        rc = fail(STM_ERR_DEVICE, "bulk-load counters");
        break;
      }
      if (bc.sort_stall) { rc = fail(STM_ERR_DEVICE, "bulk-load sort: a look-back wait gave up"); break; }
      if (bc.n_bad) { rc = fail(STM_ERR_INVALID, "bulk load holds the reserved key"); break; }
      U = bc.n_unique;
    }
    // Level sizes.
    std::vector<uint32_t> level_n;
    uint64_t L = U == 0 ? 1 : (U + fill - 1) / fill;
    level_n.push_back((uint32_t)L);
    while (level_n.back() > 1) level_n.push_back((level_n.back() + fill - 1) / fill);
    uint64_t total = 0;
    for (uint32_t c : level_n) total += c;
    if (total > t->capacity) {
      rc = fail(STM_ERR_NOMEM, "bulk load needs %llu nodes, pool holds %llu; raise max_keys",
                (unsigned long long)total, (unsigned long long)t->capacity);
      break;
    }
    uint64_t *lows_a = nullptr, *lows_b = nullptr;
    if (dalloc(&lows_a, L) != hipSuccess || dalloc(&lows_b, L) != hipSuccess) {
      if (lows_a) (void)hipFree(lows_a);
      rc = fail(STM_ERR_NOMEM, "bulk-load fence arrays");
      break;
    }
    hipError_t r2 = hipMemsetAsync(t->view.meta, 0, sizeof(NodeMeta) * t->capacity, st);
    if (r2 == hipSuccess)
      r2 = launch_build_leaves(t->view, tmp.ukey, tmp.uval, U, fill, (uint32_t)L, lows_a, st);
    uint32_t first_prev = 0, first_this = (uint32_t)L;
    uint64_t* lp = lows_a;
    uint64_t* lt = lows_b;
    for (size_t lv = 1; lv < level_n.size() && r2 == hipSuccess; ++lv) {
      r2 = launch_build_inner(t->view, lp, level_n[lv - 1], first_prev, fill, level_n[lv], first_this,
                              (uint32_t)lv, lt, st);
      first_prev = first_this;
      first_this += level_n[lv];
      std::swap(lp, lt);
    }
    if (r2 == hipSuccess) {
      TreeHeader h{};
      h.lock = 0;
      h.root = first_prev + level_n.back() - 1;
      if (level_n.size() == 1) h.root = 0;
      h.height = level_n.size();
      h.n_nodes = total;
      h.capacity = t->capacity;
      h.error = 0;
      h.n_keys = U;
      r2 = hipMemcpyAsync(t->view.hdr, &h, sizeof(h), hipMemcpyHostToDevice, st);
      if (r2 == hipSuccess)  // the key count restarts from the built tree
        r2 = hipMemsetAsync(t->view.stats + (size_t)ST_KEYS * STAT_STRIPES, 0,
                            sizeof(unsigned long long) * STAT_STRIPES, st);
      if (r2 == hipSuccess) r2 = hipStreamSynchronize(st);
      t->route_valid = false;  // the router describes the previous tree
      t->route_warm = false;
    }
    (void)hipFree(lows_a);
    (void)hipFree(lows_b);
    if (r2 != hipSuccess) rc = fail(STM_ERR_DEVICE, "bulk build: %s", hipGetErrorString(r2));
  } while (0);
  (void)hipStreamSynchronize(st);
  tmp.release();
  return rc;
}

// ------------------------------------------------------------------ batches without allocation
// stm_run_batch_device_noalloc keeps the scratch of its batch (sorted keys and positions, the
// sort's control words and histograms, fold tiles, per-group arrays, insert lists, counters)
// in this fixed device array instead of the tree's allocations.  Each GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the array's name to a
// kernel or copy: it stands for a host-side shadow; the device address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept per device.
constexpr size_t NOALLOC_BYTES = 8u << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

struct NoAllocSlot {
  std::once_flag once;
  uint8_t* ptr = nullptr;  // device address of `scratch` on this device
  int rc = STM_OK;
  char err[192] = {};
};
constexpr int NOALLOC_DEVICES = 64;
NoAllocSlot g_noalloc[NOALLOC_DEVICES];

// Lays the batch scratch out in the device's copy of `scratch` (set up on the first call on
// the device: address lookup, then the sort's control words cleared so its look-back status
// starts untagged).  Nothing is allocated.  The caller has made the tree's device current.
int noalloc_scratch(stm_tree* t, Scratch& s) {
  const int dev = t->cfg.device;
  if (dev < 0 || dev >= NOALLOC_DEVICES)
    return fail(STM_ERR_INVALID, "device %d outside the no-allocation table", dev);
  NoAllocSlot& d = g_noalloc[dev];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      d.rc = STM_ERR_DEVICE;
      snprintf(d.err, sizeof(d.err), "hipGetSymbolAddress: %s", hipGetErrorString(e));
      return;
    }
    Scratch probe;
    const size_t need = probe.carve(STM_NOALLOC_MAX_OPS, nullptr);
    if (need > NOALLOC_BYTES) {
      d.rc = STM_ERR_STATE;
      snprintf(d.err, sizeof(d.err), "layout of %zu bytes exceeds the %zu-byte array", need,
               NOALLOC_BYTES);
      return;
    }
    probe.carve(STM_NOALLOC_MAX_OPS, static_cast<uint8_t*>(p));
    e = sort_scratch_init(probe.hist);
    if (e != hipSuccess) {
      d.rc = STM_ERR_DEVICE;
      snprintf(d.err, sizeof(d.err), "sort control words: %s", hipGetErrorString(e));
      return;
    }
    d.ptr = static_cast<uint8_t*>(p);
  });
  if (d.rc != STM_OK) return fail(d.rc, "no-allocation scratch on device %d: %s", dev, d.err);
  s.carve(STM_NOALLOC_MAX_OPS, d.ptr);
  s.h_counters = t->scr.h_counters;  // pinned at stm_create
  return STM_OK;
}

}  // namespace

extern "C" {

const char* stm_last_error(void) { return g_last_error.c_str(); }

int stm_config_init(stm_config* cfg) {
  if (!cfg) return fail(STM_ERR_INVALID, "null config");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->fanout = 64;
  cfg->leaf_fill = 0;
  cfg->max_keys = 1ull << 24;
  cfg->max_batch = 1ull << 22;
  cfg->device = 0;
  cfg->lanes_per_op = 16;
  cfg->ops_per_group = 16;
  cfg->upsert_ops_per_wave = 512;
  cfg->max_attempts = 64;
  cfg->route = STM_ROUTE_DESCENT;
  cfg->timing = 0;
  cfg->retrain_every = 8;
  cfg->train_steps = 300;
  cfg->sort_tile_items = 16;
  cfg->use_graphs = 1;
  cfg->upsert_lanes = 0;  // automatic
  return STM_OK;
}

int stm_create(const stm_config* cfg_in, stm_tree** out) {
  if (!out) return fail(STM_ERR_INVALID, "null output handle");
  *out = nullptr;
  stm_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else stm_config_init(&cfg);
  if (cfg.fanout != 32 && cfg.fanout != 64 && cfg.fanout != 128)
    return fail(STM_ERR_INVALID, "fanout %u not in {32, 64, 128}", cfg.fanout);
  if (cfg.leaf_fill == 0) cfg.leaf_fill = cfg.fanout - cfg.fanout / 4;
  if (cfg.leaf_fill < 2 || cfg.leaf_fill > cfg.fanout)
    return fail(STM_ERR_INVALID, "leaf_fill %u not in [2, fanout]", cfg.leaf_fill);
  if (cfg.max_batch == 0 || cfg.max_batch > (1ull << 30))
    return fail(STM_ERR_INVALID, "max_batch %llu not in [1, 2^30]", (unsigned long long)cfg.max_batch);
  if (cfg.max_keys == 0) return fail(STM_ERR_INVALID, "max_keys must be positive");
  const uint32_t lp = cfg.lanes_per_op;
  if (!(lp == 8 || lp == 16 || lp == 32 || lp == 64) || lp > cfg.fanout)
    return fail(STM_ERR_INVALID, "lanes_per_op %u not in {8,16,32,64} or above fanout", lp);
  if (cfg.ops_per_group == 0) cfg.ops_per_group = 16;
  if (cfg.upsert_ops_per_wave == 0) cfg.upsert_ops_per_wave = 512;
  if (cfg.sort_tile_items == 0) cfg.sort_tile_items = 16;
  if (cfg.sort_tile_items != 8 && cfg.sort_tile_items != 16)
    return fail(STM_ERR_INVALID, "sort_tile_items %u not in {8, 16}", cfg.sort_tile_items);
  // Automatic: 8 lanes (8 inserts per wavefront round) while that keeps the leaf in at most 8
  // elements per lane (measured 0.186 vs 0.206 ms per batch at fanout 64), 16 above.
  if (cfg.upsert_lanes == 0) cfg.upsert_lanes = cfg.fanout <= 64 ? 8u : 16u;
  if (!(cfg.upsert_lanes == 8 || cfg.upsert_lanes == 16 || cfg.upsert_lanes == 32) ||
      cfg.upsert_lanes > cfg.fanout)
    return fail(STM_ERR_INVALID, "upsert_lanes %u not in {8, 16, 32} or above fanout", cfg.upsert_lanes);
  if (cfg.max_attempts == 0) cfg.max_attempts = 64;
  if (cfg.retrain_every == 0) cfg.retrain_every = 8;
  if (cfg.train_steps > 100000) return fail(STM_ERR_INVALID, "train_steps above 100000");
  if (cfg.route > STM_ROUTE_INDEX) return fail(STM_ERR_INVALID, "route %u unknown", cfg.route);
  const uint64_t cap = capacity_for(cfg.max_keys, cfg.fanout);
  if (cap >= NODE_NONE) return fail(STM_ERR_INVALID, "max_keys too large for 32-bit node ids");

  stm_tree* t = new stm_tree();
  t->cfg = cfg;
  t->capacity = cap;
  int rc = set_device(t);
  if (rc) {
    delete t;
    return rc;
  }
  auto bail = [&](int code) {
    stm_destroy(t);
    return code;
  };
  if (hipStreamCreate(&t->stream) != hipSuccess) return bail(fail(STM_ERR_DEVICE, "hipStreamCreate"));
  const uint32_t F = cfg.fanout;
  hipError_t e = hipSuccess;
  if ((e = dalloc(&t->view.keys, cap * F)) != hipSuccess ||
      (e = dalloc(&t->view.vals, cap * F)) != hipSuccess ||
      (e = dalloc(&t->view.meta, cap)) != hipSuccess ||
      (e = dalloc(&t->view.hdr, 1)) != hipSuccess ||
      (e = dalloc(&t->view.stats, (size_t)STATS_WORDS)) != hipSuccess)
    return bail(fail(STM_ERR_NOMEM, "node pool of %llu nodes: %s", (unsigned long long)cap,
                     hipGetErrorString(e)));
  t->view.capacity = (uint32_t)cap;
  t->view.fanout = F;
  if ((e = t->scr.alloc(cfg.max_batch, true)) != hipSuccess)
    return bail(fail(STM_ERR_NOMEM, "batch scratch for %llu ops: %s",
                     (unsigned long long)cfg.max_batch, hipGetErrorString(e)));
  if (hipMemsetAsync(t->view.meta, 0, sizeof(NodeMeta) * cap, t->stream) != hipSuccess ||
      hipMemsetAsync(t->view.stats, 0, sizeof(unsigned long long) * STATS_WORDS, t->stream) != hipSuccess)
    return bail(fail(STM_ERR_DEVICE, "pool init"));
  // Phase events only time device work: without the system-scope fence a record is a
  // timestamp, not a cache write-back between two kernels (≈10 µs each at a phase boundary).
  for (int p = 0; p < STM_NUM_PHASES; ++p) {
    if (hipEventCreateWithFlags(&t->ev_a[p], hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&t->ev_b[p], hipEventDisableSystemFence) != hipSuccess)
      return bail(fail(STM_ERR_DEVICE, "hipEventCreate"));
  }
  rc = init_empty(t);
  if (rc) return bail(rc);
  *out = t;
  return STM_OK;
}

int stm_destroy(stm_tree* t) {
  if (!t) return STM_OK;
  (void)hipSetDevice(t->cfg.device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  if (t->stream_front) (void)hipStreamSynchronize(t->stream_front);
  drop_graphs(t);
  t->scr.release();
  if (t->scr2_alloc) t->scr2.release();
  for (int k = 0; k < 2; ++k) {
    if (t->ev_front[k]) (void)hipEventDestroy(t->ev_front[k]);
    if (t->ev_back[k]) (void)hipEventDestroy(t->ev_back[k]);
  }
  if (t->ev_sync) (void)hipEventDestroy(t->ev_sync);
  if (t->stream_front) (void)hipStreamDestroy(t->stream_front);
  void* ps[] = {t->view.keys, t->view.vals, t->view.meta, t->view.hdr, t->view.stats};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (int p = 0; p < STM_NUM_PHASES; ++p) {
    if (t->ev_a[p]) (void)hipEventDestroy(t->ev_a[p]);
    if (t->ev_b[p]) (void)hipEventDestroy(t->ev_b[p]);
  }
  free_router(t);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
  return STM_OK;
}

int stm_size(stm_tree* t, uint64_t* n_keys) {
  if (!t || !n_keys) return fail(STM_ERR_INVALID, "null argument");
  int rc = set_device(t);
  if (rc) return rc;
  TreeHeader h;
  rc = read_header(t, &h);
  if (rc) return rc;
  uint64_t c[ST_NUM];
  rc = read_counters(t, c);
  if (rc) return rc;
  *n_keys = h.n_keys + c[ST_KEYS];
  return STM_OK;
}

int stm_bulk_load_device(stm_tree* t, const uint64_t* d_keys, const uint64_t* d_vals, size_t n) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n && (!d_keys || !d_vals)) return fail(STM_ERR_INVALID, "null key or value array");
  int rc = set_device(t);
  if (rc) return rc;
  uint64_t have = 0;
  rc = stm_size(t, &have);
  if (rc) return rc;
  if (have == 0) return build_from_device(t, d_keys, d_vals, n);
  // Non-empty tree: a batch of inserts, outputs discarded into the staging buffers.
  const uint64_t mb = t->cfg.max_batch;
  HIP_TRY(hipMemsetAsync(t->scr.in_ops, STM_OP_INSERT, mb, t->stream));
  for (size_t off = 0; off < n; off += mb) {
    const uint32_t m = (uint32_t)std::min<uint64_t>(mb, n - off);
    rc = run_chunk(t, d_keys + off, d_vals + off, t->scr.in_ops, m, t->scr.out_vals, t->scr.out_found);
    if (rc) return rc;
  }
  return STM_OK;
}

int stm_bulk_load(stm_tree* t, const uint64_t* keys, const uint64_t* vals, size_t n) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n && (!keys || !vals)) return fail(STM_ERR_INVALID, "null key or value array");
  int rc = set_device(t);
  if (rc) return rc;
  rc = validate_host_batch(keys, nullptr, n, false, vals);
  if (rc) return rc;
  uint64_t *dk = nullptr, *dv = nullptr;
  if (dalloc(&dk, n) != hipSuccess || dalloc(&dv, n) != hipSuccess) {
    if (dk) (void)hipFree(dk);
    return fail(STM_ERR_NOMEM, "bulk-load input of %zu keys", n);
  }
  rc = STM_OK;
  if (n) {
    if (hipMemcpyAsync(dk, keys, n * 8, hipMemcpyHostToDevice, t->stream) != hipSuccess ||
        hipMemcpyAsync(dv, vals, n * 8, hipMemcpyHostToDevice, t->stream) != hipSuccess)
      rc = fail(STM_ERR_DEVICE, "bulk-load copy");
  }
  if (!rc) rc = stm_bulk_load_device(t, dk, dv, n);
  (void)hipStreamSynchronize(t->stream);
  (void)hipFree(dk);
  (void)hipFree(dv);
  return rc;
}

int stm_run_batch_device(stm_tree* t, const uint64_t* d_keys, const uint64_t* d_vals,
                         const uint8_t* d_ops, size_t n, uint64_t* d_out_vals, uint8_t* d_out_found) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n == 0) return STM_OK;
  if (!d_keys || !d_vals || !d_ops || !d_out_vals || !d_out_found)
    return fail(STM_ERR_INVALID, "null device array (device batches always carry values)");
  int rc = set_device(t);
  if (rc) return rc;
  return run_device(t, d_keys, d_vals, d_ops, n, d_out_vals, d_out_found);
}

int stm_run_batch_device_noalloc(stm_tree* t, const uint64_t* d_keys, const uint64_t* d_vals,
                                 const uint8_t* d_ops, size_t n, uint64_t* d_out_vals,
                                 uint8_t* d_out_found) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n == 0) return STM_OK;
  if (!d_keys || !d_vals || !d_ops || !d_out_vals || !d_out_found)
    return fail(STM_ERR_INVALID, "null device array (device batches always carry values)");
  if (n > STM_NOALLOC_MAX_OPS)
    return fail(STM_ERR_INVALID, "batch of %zu ops exceeds the %u ops the no-allocation scratch "
                "holds", n, (unsigned)STM_NOALLOC_MAX_OPS);
  int rc = set_device(t);
  if (rc) return rc;
  if (t->cfg.route != STM_ROUTE_DESCENT && !t->route_alloc)
    return fail(STM_ERR_STATE, "routing buffers do not exist yet: they are allocated by the first "
                "training (stm_retrain or one stm_run_batch_device call)");
  Scratch s;
  rc = noalloc_scratch(t, s);
  if (rc) return rc;
  return run_chunk(t, d_keys, d_vals, d_ops, (uint32_t)n, d_out_vals, d_out_found, nullptr, &s);
}

int stm_run_batches_device(stm_tree* t, const uint64_t* d_keys, const uint64_t* d_vals,
                           const uint8_t* d_ops, const uint64_t* offsets, size_t n_batches,
                           uint64_t* d_out_vals, uint8_t* d_out_found) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n_batches == 0) return STM_OK;
  if (!offsets) return fail(STM_ERR_INVALID, "null offsets");
  for (size_t b = 0; b < n_batches; ++b)
    if (offsets[b + 1] < offsets[b])
      return fail(STM_ERR_INVALID, "offsets decrease at batch %zu", b);
  if (offsets[n_batches] == offsets[0]) return STM_OK;
  if (!d_keys || !d_vals || !d_ops || !d_out_vals || !d_out_found)
    return fail(STM_ERR_INVALID, "null device array (device batches always carry values)");
  int rc = set_device(t);
  if (rc) return rc;
  if (!t->cfg.use_graphs) {
    for (size_t b = 0; b < n_batches; ++b) {
      const uint64_t o = offsets[b];
      rc = run_device(t, d_keys + o, d_vals + o, d_ops + o, offsets[b + 1] - o, d_out_vals + o,
                      d_out_found + o);
      if (rc) return rc;
    }
    return STM_OK;
  }
  return run_stream(t, d_keys, d_vals, d_ops, offsets, n_batches, d_out_vals, d_out_found);
}

int stm_run_batch_scan_device(stm_tree* t, const uint64_t* d_keys, const uint64_t* d_vals,
                              const uint8_t* d_ops, size_t n, uint64_t* d_out_vals,
                              uint8_t* d_out_found, uint64_t* d_scan_keys, uint64_t* d_scan_vals,
                              size_t scan_cap) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n == 0) return STM_OK;
  if (!d_keys || !d_vals || !d_ops || !d_out_vals || !d_out_found)
    return fail(STM_ERR_INVALID, "null device array (device batches always carry values)");
  if (scan_cap && (!d_scan_keys || !d_scan_vals)) return fail(STM_ERR_INVALID, "null scan area");
  int rc = set_device(t);
  if (rc) return rc;
  const uint64_t mb = t->cfg.max_batch;
  uint64_t base = 0;
  for (size_t off = 0; off < n; off += mb) {
    const uint32_t m = (uint32_t)std::min<uint64_t>(mb, n - off);
    ScanOut so{d_scan_keys ? d_scan_keys + base : nullptr, d_scan_vals ? d_scan_vals + base : nullptr,
               scan_cap - base, 0};
    rc = run_chunk(t, d_keys + off, d_vals + off, d_ops + off, m, d_out_vals + off,
                   d_out_found + off, &so);
    if (rc) return rc;
    base += std::min<uint64_t>(so.used, scan_cap - base);
  }
  return STM_OK;
}

int stm_run_batch_scan(stm_tree* t, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                       size_t n, uint64_t* out_vals, uint8_t* out_found, uint64_t* scan_keys,
                       uint64_t* scan_vals, size_t scan_cap) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n == 0) return STM_OK;
  if (!keys || !vals || !ops || !out_vals || !out_found) return fail(STM_ERR_INVALID, "null array");
  int rc = set_device(t);
  if (rc) return rc;
  rc = validate_host_batch(keys, ops, n, true, vals);
  if (rc) return rc;
  // Scan lengths: checked here, and each chunk's output area sized from them.
  const uint64_t mb = t->cfg.max_batch;
  std::vector<uint64_t> chunk_pairs;
  uint64_t total = 0;
  for (size_t off = 0; off < n; off += mb) {
    const size_t m = (size_t)std::min<uint64_t>(mb, n - off);
    uint64_t c = 0;
    for (size_t i = off; i < off + m; ++i) {
      if (ops[i] != STM_OP_RANGE) continue;
      if (vals[i] < 1 || vals[i] > STM_SCAN_MAX_LEN)
        return fail(STM_ERR_INVALID, "scan op %zu has length %llu, not in [1, %u]", i,
                    (unsigned long long)vals[i], STM_SCAN_MAX_LEN);
      c += vals[i];
    }
    chunk_pairs.push_back(c);
    total += c;
  }
  if (total > scan_cap)
    return fail(STM_ERR_INVALID, "scans need %llu pairs, the area holds %zu",
                (unsigned long long)total, scan_cap);
  if (total && (!scan_keys || !scan_vals)) return fail(STM_ERR_INVALID, "null scan area");
  const uint64_t most = *std::max_element(chunk_pairs.begin(), chunk_pairs.end());
  uint64_t *dsk = nullptr, *dsv = nullptr;
  if (dalloc(&dsk, most) != hipSuccess || dalloc(&dsv, most) != hipSuccess) {
    if (dsk) (void)hipFree(dsk);
    return fail(STM_ERR_NOMEM, "scan output area of %llu pairs", (unsigned long long)most);
  }
  Scratch& s = t->scr;
  uint64_t base = 0;
  size_t ci = 0;
  for (size_t off = 0; off < n && rc == STM_OK; off += mb, ++ci) {
    const size_t m = (size_t)std::min<uint64_t>(mb, n - off);
    hipError_t e = hipMemcpyAsync(s.in_keys, keys + off, m * 8, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.in_vals, vals + off, m * 8, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.in_ops, ops + off, m, hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) {
      rc = fail(STM_ERR_DEVICE, "batch copy: %s", hipGetErrorString(e));
      break;
    }
    ScanOut so{dsk, dsv, std::max<uint64_t>(most, 1), 0};
    rc = run_chunk(t, s.in_keys, s.in_vals, s.in_ops, (uint32_t)m, s.out_vals, s.out_found, &so);
    if (rc) break;
    const uint64_t c = chunk_pairs[ci];
    e = hipMemcpyAsync(out_vals + off, s.out_vals, m * 8, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out_found + off, s.out_found, m, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && c) e = hipMemcpyAsync(scan_keys + base, dsk, c * 8, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && c) e = hipMemcpyAsync(scan_vals + base, dsv, c * 8, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) rc = fail(STM_ERR_DEVICE, "result copy: %s", hipGetErrorString(e));
    base += c;
  }
  (void)hipStreamSynchronize(t->stream);
  (void)hipFree(dsk);
  (void)hipFree(dsv);
  return rc;
}

int stm_sort_pairs_device(int device, const uint64_t* d_keys, size_t n, uint64_t* d_keys_out,
                          uint32_t* d_idx_out) {
  if (n == 0) return STM_OK;
  if (!d_keys || !d_keys_out || !d_idx_out) return fail(STM_ERR_INVALID, "null device array");
  if (n > (1ull << 30)) return fail(STM_ERR_INVALID, "sort of %zu keys exceeds 2^30", n);
  HIP_TRY(hipSetDevice(device));
  uint64_t* ktmp = nullptr;
  uint32_t *itmp = nullptr, *words = nullptr;
  int rc = STM_OK;
  if (dalloc(&ktmp, n) != hipSuccess || dalloc(&itmp, n) != hipSuccess ||
      dalloc(&words, sort_hist_words(n)) != hipSuccess) {
    rc = fail(STM_ERR_NOMEM, "sort scratch for %zu keys", n);
  } else {
    hipError_t e = sort_scratch_init(words);
    if (e == hipSuccess)
      e = radix_sort_pairs(d_keys, ktmp, d_keys_out, itmp, d_idx_out, (uint32_t)n, 0, 64, words,
                           nullptr);
    uint32_t stalled = 0;
    if (e == hipSuccess)
      e = hipMemcpy(&stalled, sort_stall_word(words), sizeof(stalled), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(STM_ERR_DEVICE, "sort: %s", hipGetErrorString(e));
    else if (stalled) rc = fail(STM_ERR_DEVICE, "sort: a bounded wait between workgroups gave up");
  }
  (void)hipDeviceSynchronize();
  void* ps[] = {ktmp, itmp, words};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  return rc;
}

int stm_capture_graphs(stm_tree* t) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  int rc = set_device(t);
  if (rc) return rc;
  rc = maybe_retrain(t);  // learned routing: the router's buffers exist before capture
  if (rc) return rc;
  if (t->graphs_exact != (t->exact_left > 0)) {
    drop_graphs(t);
    t->graphs_exact = t->exact_left > 0;
  }
  const int last = class_of(t, (uint32_t)t->cfg.max_batch);
  for (int c = 0; c <= last; ++c) {
    rc = capture_class(t, c);
    if (rc) return rc;
  }
  return STM_OK;
}

int stm_run_batch(stm_tree* t, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                  size_t n, uint64_t* out_vals, uint8_t* out_found) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  if (n == 0) return STM_OK;
  if (!keys || !ops || !out_vals || !out_found) return fail(STM_ERR_INVALID, "null array");
  int rc = set_device(t);
  if (rc) return rc;
  rc = validate_host_batch(keys, ops, n, true, vals);
  if (rc) return rc;
  Scratch& s = t->scr;
  const uint64_t mb = t->cfg.max_batch;
  for (size_t off = 0; off < n; off += mb) {
    const size_t m = (size_t)std::min<uint64_t>(mb, n - off);
    HIP_TRY(hipMemcpyAsync(s.in_keys, keys + off, m * 8, hipMemcpyHostToDevice, t->stream));
    if (vals) HIP_TRY(hipMemcpyAsync(s.in_vals, vals + off, m * 8, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipMemcpyAsync(s.in_ops, ops + off, m, hipMemcpyHostToDevice, t->stream));
    rc = run_chunk(t, s.in_keys, vals ? s.in_vals : nullptr, s.in_ops, (uint32_t)m, s.out_vals,
                   s.out_found);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out_vals + off, s.out_vals, m * 8, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipMemcpyAsync(out_found + off, s.out_found, m, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
  }
  return STM_OK;
}

int stm_dump(stm_tree* t, uint64_t* keys, uint64_t* vals, size_t cap, size_t* n) {
  if (!t || !n) return fail(STM_ERR_INVALID, "null argument");
  int rc = set_device(t);
  if (rc) return rc;
  TreeHeader h;
  rc = read_header(t, &h);
  if (rc) return rc;
  const uint32_t nn = (uint32_t)std::min<uint64_t>(h.n_nodes, t->capacity);
  hipStream_t st = t->stream;
  uint64_t *lows = nullptr, *slows = nullptr, *slows_tmp = nullptr, *offs = nullptr;
  uint32_t *ids = nullptr, *order = nullptr, *order_tmp = nullptr, *sorted_ids = nullptr,
           *counts = nullptr, *counter = nullptr, *hist = nullptr;
  uint64_t *ok = nullptr, *ov = nullptr;
  auto cleanup = [&]() {
    void* ps[] = {lows, slows, slows_tmp, offs, ids, order, order_tmp, sorted_ids, counts, counter,
                  hist, ok, ov};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  auto err = [&](int code, const char* what) {
    (void)hipStreamSynchronize(st);
    cleanup();
    return fail(code, "dump: %s", what);
  };
  if (dalloc(&lows, nn) || dalloc(&slows, nn) || dalloc(&slows_tmp, nn) || dalloc(&ids, nn) ||
      dalloc(&order, nn) || dalloc(&order_tmp, nn) || dalloc(&sorted_ids, nn) ||
      dalloc(&counts, nn) || dalloc(&counter, 1) || dalloc(&offs, (size_t)nn + 1) ||
      dalloc(&hist, sort_hist_words(nn)))
    return err(STM_ERR_NOMEM, "scratch");
  if (sort_scratch_init(hist) != hipSuccess) return err(STM_ERR_DEVICE, "sort scratch");
  if (hipMemsetAsync(counter, 0, 4, st) != hipSuccess) return err(STM_ERR_DEVICE, "memset");
  if (launch_collect_leaves(t->view, nn, lows, ids, counter, st) != hipSuccess)
    return err(STM_ERR_DEVICE, "collect");
  uint32_t n_leaves = 0;
  if (hipMemcpyAsync(&n_leaves, counter, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return err(STM_ERR_DEVICE, "leaf count");
  if (radix_sort_pairs(lows, slows_tmp, slows, order_tmp, order, n_leaves, 0, 64, hist, st) != hipSuccess ||
      launch_leaf_counts(t->view, order, ids, n_leaves, sorted_ids, counts, st) != hipSuccess)
    return err(STM_ERR_DEVICE, "leaf order");
  uint32_t stalled = 0;
  if (hipMemcpyAsync(&stalled, sort_stall_word(hist), 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return err(STM_ERR_DEVICE, "leaf order");
  if (stalled) return err(STM_ERR_DEVICE, "leaf order: a sort look-back wait gave up");
  std::vector<uint32_t> hc(n_leaves);
  if (n_leaves && (hipMemcpyAsync(hc.data(), counts, 4ull * n_leaves, hipMemcpyDeviceToHost, st) != hipSuccess ||
                   hipStreamSynchronize(st) != hipSuccess))
    return err(STM_ERR_DEVICE, "counts");
  std::vector<uint64_t> ho(n_leaves + 1, 0);
  for (uint32_t r = 0; r < n_leaves; ++r) ho[r + 1] = ho[r] + hc[r];
  const uint64_t total = ho[n_leaves];
  *n = total;
  if (!keys || !vals || cap < total) {
    (void)hipStreamSynchronize(st);
    cleanup();
    if (!keys && !vals && cap == 0) return STM_OK;  // size query
    return fail(STM_ERR_INVALID, "dump needs room for %llu pairs, got %zu", (unsigned long long)total, cap);
  }
  if (dalloc(&ok, total) || dalloc(&ov, total)) return err(STM_ERR_NOMEM, "output");
  if (hipMemcpyAsync(offs, ho.data(), 8ull * (n_leaves + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
      launch_dump_copy(t->view, sorted_ids, offs, n_leaves, ok, ov, st) != hipSuccess)
    return err(STM_ERR_DEVICE, "copy");
  if (total && (hipMemcpyAsync(keys, ok, total * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(vals, ov, total * 8, hipMemcpyDeviceToHost, st) != hipSuccess))
    return err(STM_ERR_DEVICE, "download");
  if (hipStreamSynchronize(st) != hipSuccess) return err(STM_ERR_DEVICE, "sync");
  cleanup();
  return STM_OK;
}

int stm_check(stm_tree* t, uint64_t* bad) {
  if (!t || !bad) return fail(STM_ERR_INVALID, "null argument");
  int rc = set_device(t);
  if (rc) return rc;
  TreeHeader h;
  rc = read_header(t, &h);
  if (rc) return rc;
  unsigned long long* d = nullptr;
  HIP_TRY(dalloc(&d, 1));
  unsigned long long v = 0;
  hipError_t e = hipMemsetAsync(d, 0, 8, t->stream);
  const uint32_t nn = (uint32_t)std::min<uint64_t>(h.n_nodes, t->capacity);
  if (e == hipSuccess) e = launch_check(t->view, nn, d, t->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, t->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(STM_ERR_DEVICE, "check: %s", hipGetErrorString(e));
  // Root and height must describe a node of the right level.
  NodeMeta rm;
  HIP_TRY(hipMemcpy(&rm, t->view.meta + h.root, sizeof(rm), hipMemcpyDeviceToHost));
  if (meta_level(rm.info) + 1 != h.height || rm.low != 0 || rm.high != KEY_EMPTY) v++;
  *bad = v;
  return STM_OK;
}

int stm_get_stats(stm_tree* t, stm_stats* out) {
  if (!t || !out) return fail(STM_ERR_INVALID, "null argument");
  int rc = set_device(t);
  if (rc) return rc;
  TreeHeader h;
  rc = read_header(t, &h);
  if (rc) return rc;
  uint64_t c[ST_NUM];
  rc = read_counters(t, c);
  if (rc) return rc;
  *out = t->st;
  out->n_keys = h.n_keys + c[ST_KEYS];
  out->n_nodes = std::min<uint64_t>(h.n_nodes, t->capacity);
  out->capacity_nodes = t->capacity;
  out->height = h.height;
  out->upserts = c[ST_UPSERTS];
  out->commits = c[ST_COMMITS];
  out->aborts = c[ST_ABORTS];
  out->leaf_splits = c[ST_LEAF_SPLITS];
  out->inner_splits = c[ST_INNER_SPLITS];
  out->root_splits = c[ST_ROOT_SPLITS];
  out->deferred = c[ST_DEFERRED];
  out->route_hits = c[ST_ROUTE_HITS];
  out->route_misses = c[ST_ROUTE_MISSES];
  out->node_reads = c[ST_NODE_READS];
  out->leaf_steps = c[ST_LEAF_STEPS];
  out->sort_fallbacks = t->sort_fallbacks;
  return STM_OK;
}

int stm_retrain(stm_tree* t, uint32_t train_steps) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  int rc = set_device(t);
  if (rc) return rc;
  rc = train_router(t, train_steps, false);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(t->stream));
  phase_collect(t);
  return STM_OK;
}

int stm_route_info(stm_tree* t, uint64_t* n_leaves, uint64_t* window_hits, uint64_t* window_misses) {
  if (!t || !n_leaves || !window_hits || !window_misses) return fail(STM_ERR_INVALID, "null argument");
  if (!t->route_valid) return fail(STM_ERR_STATE, "no router has been trained (stm_retrain)");
  int rc = set_device(t);
  if (rc) return rc;
  uint64_t* d = nullptr;
  HIP_TRY(dalloc(&d, 2));
  uint64_t hm[2] = {0, 0};
  hipError_t e = hipMemsetAsync(d, 0, 16, t->stream);
  if (e == hipSuccess) e = route_evaluate(t->route, d, (uint32_t)t->capacity, t->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hm, d, 16, hipMemcpyDeviceToHost, t->stream);
  RouteHeader rh{};
  if (e == hipSuccess) e = hipMemcpyAsync(&rh, t->route.hdr, sizeof(rh), hipMemcpyDeviceToHost, t->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(STM_ERR_DEVICE, "route info: %s", hipGetErrorString(e));
  *n_leaves = rh.L;
  *window_hits = hm[0];
  *window_misses = hm[1];
  return STM_OK;
}

int stm_route_directory(stm_tree* t, uint64_t* fences, uint32_t* leaf_ids, uint64_t max,
                        uint64_t* n_leaves) {
  if (!t || !n_leaves || (!fences != !leaf_ids)) return fail(STM_ERR_INVALID, "null argument");
  if (!t->route_valid) return fail(STM_ERR_STATE, "no router has been trained (stm_retrain)");
  int rc = set_device(t);
  if (rc) return rc;
  RouteHeader rh{};
  HIP_TRY(hipMemcpyAsync(&rh, t->route.hdr, sizeof(rh), hipMemcpyDeviceToHost, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  *n_leaves = rh.L;
  if (!fences) return STM_OK;  // size query
  if (rh.L > max) return fail(STM_ERR_INVALID, "directory holds %u leaves, room for %llu", rh.L,
                              (unsigned long long)max);
  if (rh.L) {
    HIP_TRY(hipMemcpyAsync(fences, t->route.fences, sizeof(uint64_t) * rh.L, hipMemcpyDeviceToHost,
                           t->stream));
    HIP_TRY(hipMemcpyAsync(leaf_ids, t->route.lids, sizeof(uint32_t) * rh.L, hipMemcpyDeviceToHost,
                           t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
  }
  return STM_OK;
}

int stm_mlp_gradients(const float* x, const float* tg, size_t n, const float* params49,
                      float* grad49, float* loss) {
  if (!params49 || !grad49 || !loss || (n && (!x || !tg))) return fail(STM_ERR_INVALID, "null argument");
  if (n > (1u << 24)) return fail(STM_ERR_INVALID, "at most 2^24 pairs");
  float *dx = nullptr, *dt = nullptr, *dp = nullptr, *dg = nullptr, *dl = nullptr;
  int rc = STM_OK;
  if (dalloc(&dx, n) || dalloc(&dt, n) || dalloc(&dp, MLP_PARAMS) || dalloc(&dg, MLP_PARAMS) ||
      dalloc(&dl, 1)) {
    rc = fail(STM_ERR_NOMEM, "gradient buffers");
  } else {
    hipError_t e = hipSuccess;
    if (n) e = hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(dt, tg, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, params49, MLP_PARAMS * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_grad(dx, dt, (uint32_t)n, dp, dg, dl, nullptr);
    if (e == hipSuccess) e = hipMemcpy(grad49, dg, MLP_PARAMS * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss, dl, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(STM_ERR_DEVICE, "mlp gradients: %s", hipGetErrorString(e));
  }
  void* ps[] = {dx, dt, dp, dg, dl};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  return rc;
}

int stm_reset_stats(stm_tree* t) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  int rc = set_device(t);
  if (rc) return rc;
  t->st = stm_stats{};
  t->sort_fallbacks = 0;
  HIP_TRY(hipMemsetAsync(t->view.stats, 0, sizeof(unsigned long long) * STATS_RESET_WORDS, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  return STM_OK;
}

int stm_get_stream(stm_tree* t, void** stream) {
  if (!t || !stream) return fail(STM_ERR_INVALID, "null argument");
  *stream = (void*)t->stream;
  return STM_OK;
}

int stm_device_alloc(int device, size_t bytes, void** ptr) {
  if (!ptr) return fail(STM_ERR_INVALID, "null argument");
  *ptr = nullptr;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return STM_OK;
}

int stm_device_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return STM_OK;
}

int stm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  if (!bytes) return STM_OK;
  if (!dst || !src) return fail(STM_ERR_INVALID, "null argument");
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return STM_OK;
}

int stm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  if (!bytes) return STM_OK;
  if (!dst || !src) return fail(STM_ERR_INVALID, "null argument");
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return STM_OK;
}

int stm_device_copy(void* dst, const void* src, size_t bytes) {
  if (!bytes) return STM_OK;
  if (!dst || !src) return fail(STM_ERR_INVALID, "null argument");
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return fail(STM_ERR_INVALID, "device copy needs 16-byte aligned pointers and size");
  const uint64_t n16 = bytes / 16;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, COPY_MAX_BLOCKS);
  k_copy16<<<blocks, 256, 0, nullptr>>>((uint4*)dst, (const uint4*)src, n16);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return STM_OK;
}

int stm_sync(stm_tree* t) {
  if (!t) return fail(STM_ERR_INVALID, "null tree");
  HIP_TRY(hipStreamSynchronize(t->stream));
  return STM_OK;
}

}  // extern "C"
