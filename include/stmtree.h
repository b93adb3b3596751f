/*
 * stmtree.h — C ABI of the MI355X-native transactional B+tree batch path.
 *
 * The library keeps an ordered u64 -> u64 map resident in HBM and runs batches of
 * mixed point lookups and inserts whose results equal a sequential replay of the
 * batch in input order (the reference's sort + per-key fold linearisation).
 *
 * Every entry point returns an int status (STM_OK or an STM_ERR_* code) and never
 * exits the process; stm_last_error() gives the message of the last failure on the
 * calling thread.  A tree handle is not thread-safe: one host thread drives it.
 *
 * Reference interface each entry point replaces (paths relative to the reference repo):
 *   GPU_BTree::prepareGPU          src/tree.h:32, src/gpu-stm-tree-v38/bpt-gpustm.cpp:14-28   -> stm_create
 *   GPU_BTree::~GPU_BTree          src/tree.h:30, src/gpu-stm-tree-v38/bpt-gpustm.h:201         -> stm_destroy
 *   first "-i" insert-only file    README.md:9, src/main.cpp:25-29                             -> stm_bulk_load
 *   prepareInput + gpu_work        src/tree.h:33,35, src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-174 -> stm_run_batch
 *   test() answer un-permute       src/gpu-stm-tree-v38/bpt-gpustm.cpp:185-192                  -> out_vals/out_found
 *   Time_Measure::getGPUTime       src/time_measure.cuh:30-42                                   -> stm_get_stats
 *   ABORT_COUNT counters           src/gpu-stm-tree-v38/bpt-gpustm.cpp:142-164                  -> stm_get_stats
 *   CUDA_ERROR_HANDLER (exit)      src/cuda_utils.h:3-12                                        -> status codes
 */
#ifndef STMTREE_H
#define STMTREE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define STM_OK 0
#define STM_ERR_INVALID 1 /* bad argument (null pointer, key == STM_KEY_RESERVED, bad op code, size) */
#define STM_ERR_NOMEM 2   /* device allocation failed or the node pool is exhausted */
#define STM_ERR_DEVICE 3  /* a HIP runtime call or a kernel failed */
#define STM_ERR_STATE 4   /* call not valid in the tree's current state */

/* Op codes: same values as the reference's mission_t (src/input-helper.h:14-17).
 * INSERT is an upsert that returns the previous value; GET returns the value.
 * As in the reference's merge (src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:34), every
 * op code other than INSERT is executed with GET semantics by stm_run_batch. */
#define STM_OP_INSERT 0
#define STM_OP_GET 1
#define STM_OP_DELETE 2
#define STM_OP_RANGE 3

/* The one key value that cannot be stored (it is the empty-slot sentinel). */
#define STM_KEY_RESERVED 0xffffffffffffffffull

/* Lookup routing methods (stm_config.route). */
#define STM_ROUTE_DESCENT 0 /* root-to-leaf descent (the plain method) */
#define STM_ROUTE_LEARNED 1 /* learned leaf-window prediction, verified, with descent fallback */
#define STM_ROUTE_INDEX 2   /* static leaf index: the same leaf directory searched without a model
                               (the non-learned comparison; reference analogue: the rebuilt index
                               of src/gpu-stm-tree-v35/bpt-gpustm-kernel.cu:17-231) */

typedef struct stm_config {
    uint32_t fanout;      /* slots per node: 32, 64 or 128 (default 64) */
    uint32_t leaf_fill;   /* keys per node written by bulk load (0 = 3/4 of fanout) */
    uint64_t max_keys;    /* node-pool sizing: keys the tree may hold (default 2^27) */
    uint64_t max_batch;   /* ops per device pass; longer batches are chunked (default 2^22) */
    int32_t device;       /* HIP device ordinal (default 0) */
    uint32_t lanes_per_op;/* lookup kernel lanes per op: 8, 16, 32 or 64 (default 16) */
    uint32_t ops_per_group;/* sorted ops walked by one lane group (default 16) */
    uint32_t upsert_ops_per_wave; /* batch ops per upsert wavefront: the upsert grid is
                                     max_batch / this many wavefronts, at most 8192 (default 512) */
    uint32_t max_attempts;/* STM attempts before an op is deferred to the serial pass (default 64) */
    uint32_t route;       /* STM_ROUTE_* */
    uint32_t timing;      /* 1 = record HIP events around every phase (stm_get_stats) */
    uint32_t retrain_every; /* learned routing: rebuild + retrain after this many batches (default 8) */
    uint32_t train_steps; /* learned routing: Adam steps per segment MLP (default 300) */
    uint32_t sort_tile_items; /* radix sort elements per thread: 8 or 16 (tiles of 2048 or 4096
                                 keys; default 16) */
    uint32_t use_graphs;  /* stm_run_batches_device: 1 = replay one captured hipGraph per batch
                             size class (default 1), 0 = launch every kernel from the host */
    uint32_t upsert_lanes;/* lanes per insert on the upsert kernel's lane-group fast path:
                             8, 16 or 32, at most fanout (64 / this many inserts per
                             wavefront at once); 0 = automatic (default): 8 up to fanout 64,
                             16 above */
    uint32_t retrain_splits; /* learned routing / static index: also rebuild the directory (and
                                retrain) before a batch once this many leaves have split since
                                the last retrain; 0 = automatic (default): a 32nd of the leaves
                                in the directory.  Counted per synchronous batch; a batch stream
                                counts its splits when it ends */
} stm_config;

/* Per-phase timing slots of stm_stats.phase_ms / phase_launches. */
#define STM_PHASE_SORT 0
#define STM_PHASE_MERGE 1
#define STM_PHASE_LOOKUP 2
#define STM_PHASE_UPSERT 3
#define STM_PHASE_DEFERRED 4
#define STM_PHASE_ROUTE 5
#define STM_PHASE_TRAIN 6
#define STM_PHASE_SCAN 7
#define STM_NUM_PHASES 8

typedef struct stm_stats {
    uint64_t n_keys;          /* keys in the tree */
    uint64_t n_nodes;         /* nodes allocated from the pool */
    uint64_t capacity_nodes;  /* node-pool capacity */
    uint64_t height;          /* levels (1 = root is a leaf) */
    uint64_t batches;         /* device passes run */
    uint64_t ops;             /* ops submitted */
    uint64_t unique_ops;      /* ops after the per-key fold */
    uint64_t upserts;         /* merged INSERT ops executed as transactions */
    uint64_t commits;         /* committed transactions (upserts + preemptive splits) */
    uint64_t aborts;          /* aborted transaction attempts */
    uint64_t leaf_splits;
    uint64_t inner_splits;
    uint64_t root_splits;
    uint64_t deferred;        /* ops that exhausted max_attempts and ran in the serial pass */
    uint64_t route_hits;      /* learned routing: leaf found inside the predicted window */
    uint64_t route_misses;    /* learned routing: fell back to descent */
    double phase_ms[STM_NUM_PHASES];
    uint64_t phase_launches[STM_NUM_PHASES];
    uint64_t node_reads;      /* tree nodes read by lookups and transactions (the reference's
                                 ABORT_COUNT traversal steps, src/gpu-stm-tree-v38/bpt-gpustm.cpp:142-160) */
    uint64_t leaf_steps;      /* right-sibling steps taken along the leaf level (learned and index
                                 routing after leaf splits; the reference's leaf traversal steps,
                                 src/gpu-stm-tree-v38/bpt-gpustm.cpp:142-160) */
    uint64_t sort_fallbacks;  /* batches above 2^18 ops whose sort of normalized keys met a long
                                 run of close, distinct keys (hot keys with neighbours) and re-sorted
                                 with the exact passes; the next 16 batches then sort exactly */
} stm_stats;

typedef struct stm_tree stm_tree;

/* Fill *cfg with defaults. */
int stm_config_init(stm_config* cfg);

/* Create an empty tree on cfg->device (cfg may be NULL for defaults). */
int stm_create(const stm_config* cfg, stm_tree** out);

/* Free every device and host resource of the tree. NULL is accepted. */
int stm_destroy(stm_tree* tree);

/* Load n (key, value) pairs from host memory.  On an empty tree the tree is built
 * bottom-up; on a non-empty tree this is a batch of n inserts.  Duplicate keys keep
 * the value of the last occurrence (sequential insert semantics). */
int stm_bulk_load(stm_tree* tree, const uint64_t* keys, const uint64_t* vals, size_t n);

/* Same as stm_bulk_load with device-resident inputs. */
int stm_bulk_load_device(stm_tree* tree, const uint64_t* d_keys, const uint64_t* d_vals, size_t n);

/* Run one batch of n ops held in host memory.  vals[i] is read only for inserts
 * (vals may be NULL when the batch has no inserts).  For every op i:
 *   out_found[i] = 1 and out_vals[i] = value  if the key was present before op i,
 *   out_found[i] = 0 and out_vals[i] = 0      otherwise,
 * where "before op i" means after every earlier op of the batch, in input order. */
int stm_run_batch(stm_tree* tree, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                  size_t n, uint64_t* out_vals, uint8_t* out_found);

/* Same with device-resident inputs and outputs (d_vals is required).  Work is queued on
 * the tree's stream (stm_get_stream); the call returns after the batch has completed.
 * Batches of at most 2^18 ops are sorted in a single sweep whose waits between workgroups are
 * bounded; a sort that gives up such a wait fails the call with STM_ERR_DEVICE, and that batch
 * then leaves the tree unchanged (its lookups, scans and inserts do not run). */
int stm_run_batch_device(stm_tree* tree, const uint64_t* d_keys, const uint64_t* d_vals,
                         const uint8_t* d_ops, size_t n, uint64_t* d_out_vals, uint8_t* d_out_found);

/* Largest batch of one stm_run_batch_device_noalloc call. */
#define STM_NOALLOC_MAX_OPS 65536

/* stm_run_batch_device for callers that cannot allocate memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No counterpart in the reference.
 * Same arguments, results and errors as stm_run_batch_device for n <= STM_NOALLOC_MAX_OPS
 * (larger n: STM_ERR_INVALID).  The batch's scratch (sorted keys and positions, sort control
 * words and histograms, fold tiles, per-key groups, insert lists, counters: about 7 MiB) lives
 * in one fixed device array per GPU, which the library holds until the process ends and sets
 * up on the first call on each device; the call allocates nothing (no hipMalloc, hipHostMalloc
 * or hipFree).  With learned or index routing the router's buffers must exist already (an
 * earlier stm_run_batch_device or stm_retrain allocates them), else STM_ERR_STATE.  Every call
 * on one GPU uses that GPU's one array, which keeps the last call's contents: calls that
 * overlap in time on one GPU (two threads, or two trees on one device) race and must not be
 * made. */
int stm_run_batch_device_noalloc(stm_tree* tree, const uint64_t* d_keys, const uint64_t* d_vals,
                                 const uint8_t* d_ops, size_t n, uint64_t* d_out_vals,
                                 uint8_t* d_out_found);

/* Longest range scan (pairs). */
#define STM_SCAN_MAX_LEN 65536

/* Run one batch in which op 3 (STM_OP_RANGE) is a range scan: SCAN(keys[i], L = vals[i]),
 * 1 <= L <= STM_SCAN_MAX_LEN, returns the first L (key, value) pairs with key >= keys[i] in
 * the state before op i (the tree plus the batch's earlier inserts), in ascending key order.
 * Scan op i writes its pairs to scan_keys / scan_vals starting at off_i = the sum of the L of
 * the scan ops before it in the batch (so the area needs the sum of all L, scan_cap >= that);
 * out_vals[i] = pairs written (< L only at the end of the key space), out_found[i] = (pairs >
 * 0); the rest of its L slots are left unwritten.  Other ops as stm_run_batch (DELETE as GET).
 * A scan with a bad length or an overflowing area fails the call with STM_ERR_INVALID after
 * the batch's other ops have run.  No reference counterpart: the reference declares RANGE but
 * never implements it (src/tree.h:37-39, src/main.cpp:30-34). */
int stm_run_batch_scan(stm_tree* tree, const uint64_t* keys, const uint64_t* vals,
                       const uint8_t* ops, size_t n, uint64_t* out_vals, uint8_t* out_found,
                       uint64_t* scan_keys, uint64_t* scan_vals, size_t scan_cap);
/* Same with device-resident inputs, outputs and scan area. */
int stm_run_batch_scan_device(stm_tree* tree, const uint64_t* d_keys, const uint64_t* d_vals,
                              const uint8_t* d_ops, size_t n, uint64_t* d_out_vals,
                              uint8_t* d_out_found, uint64_t* d_scan_keys, uint64_t* d_scan_vals,
                              size_t scan_cap);

/* Run a stream of batches stored back to back in device memory: batch b is ops
 * [offsets[b], offsets[b+1]) of d_keys / d_vals / d_ops and is answered into the same range of
 * d_out_vals / d_out_found.  offsets is a host array of n_batches + 1 non-decreasing entries;
 * a batch longer than max_batch is chunked as in stm_run_batch_device.  The results equal
 * n_batches calls of stm_run_batch_device, in order.  With use_graphs the launch sequence of a
 * batch is captured once per size class (powers of 4 up to max_batch) as a hipGraph that reads
 * its batch from a device-side descriptor table, and replayed: the host reads nothing back
 * until the stream has run (a reserved key or a full node pool is reported at the end, the
 * tree then holds every batch up to and including the failing one; the outputs of the batches
 * after it are unspecified).  The graphs come in two halves per batch on two scratch sets, so
 * batch b + 1's sort and fold run while batch b's lookups and upserts do (a second batch
 * scratch and a second stream are made on first use).
 * Replaces the batch loop of GPU_STM_Tree::gpu_work_main (src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-135:
 * transferData, launchKernelBefore, launchKernel, launchKernelAfter, transferAns per batch). */
int stm_run_batches_device(stm_tree* tree, const uint64_t* d_keys, const uint64_t* d_vals,
                           const uint8_t* d_ops, const uint64_t* offsets, size_t n_batches,
                           uint64_t* d_out_vals, uint8_t* d_out_found);

/* Captures the graphs of every batch size class now (stm_run_batches_device captures a class
 * the first time a batch of that size arrives), so a latency-sensitive caller pays no capture
 * inside its stream.  No reference counterpart (the reference launches every kernel from the
 * host, src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:264-332). */
int stm_capture_graphs(stm_tree* tree);

/* Stable ascending sort of n device-resident u64 keys on `device`: d_keys_out receives the
 * keys in order, d_idx_out the input position of each (ties keep input order).  The batch
 * path's own sort, exposed for callers and tests; it allocates its scratch per call.
 * Replaces cub::DeviceRadixSort::SortPairs of (key, position)
 * (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:293). */
int stm_sort_pairs_device(int device, const uint64_t* d_keys, size_t n, uint64_t* d_keys_out,
                          uint32_t* d_idx_out);

/* Number of keys in the tree. */
int stm_size(stm_tree* tree, uint64_t* n_keys);

/* In-order dump of all (key, value) pairs into host buffers of capacity cap.
 * *n receives the number of keys; STM_ERR_INVALID if cap < *n. */
int stm_dump(stm_tree* tree, uint64_t* keys, uint64_t* vals, size_t cap, size_t* n);

/* Structural check of the device tree: inner nodes in key order, leaf entries distinct (a
 * leaf's entries are unordered), keys within their fences, free slots empty, levels, sibling
 * chain, child ranges, all locks released.  *bad receives the number of violations. */
int stm_check(stm_tree* tree, uint64_t* bad);

/* Counters and per-phase device times accumulated since create or stm_reset_stats. */
int stm_get_stats(stm_tree* tree, stm_stats* out);
int stm_reset_stats(stm_tree* tree);

/* The hipStream_t every kernel of this tree runs on. */
int stm_get_stream(stm_tree* tree, void** stream);

/* Learned routing (no reference counterpart; the reference always descends from the root).
 * stm_retrain rebuilds the leaf directory from the current tree and trains the router on
 * the device now (train_steps = 0 keeps the untrained, identity-like segment models).
 * stm_route_info reports the directory size and, for every leaf low fence, whether the
 * current model's +/-2 window contains it (hits) or not (misses). */
int stm_retrain(stm_tree* tree, uint32_t train_steps);
int stm_route_info(stm_tree* tree, uint64_t* n_leaves, uint64_t* window_hits, uint64_t* window_misses);
/* Copies the leaf directory (low fence and leaf id of every leaf, in key order) to host
 * arrays of max entries; *n_leaves receives its size.  STM_ERR_INVALID when max is smaller;
 * with both arrays null only the size is returned. */
int stm_route_directory(stm_tree* tree, uint64_t* fences, uint32_t* leaf_ids, uint64_t max,
                        uint64_t* n_leaves);

/* One forward/backward pass of the router's segment MLP (1 -> 16 -> 1, ReLU, fp32) on the
 * device, for checking the training kernels: loss = mean of 0.5 * (mlp(x_i) - t_i)^2 and its
 * gradient with respect to the 49 parameters (w1[16], b1[16], w2[16], b2), host arrays. */
int stm_mlp_gradients(const float* x, const float* t, size_t n, const float* params49,
                      float* grad49, float* loss);

/* ---------------------------------------------------------------- serving loop
 * Continuous batching over one device-resident tree (no reference counterpart; it generalises
 * the fixed batch loop of GPU_STM_Tree::gpu_work_main, src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-135).
 * Requests (host arrays, mixed GET / INSERT / SCAN ops as in stm_run_batch_scan) are admitted
 * FIFO into the batch being formed while it has room in its slot of the HBM budget; batches run
 * one after another on the tree while the next forms and the previous drains, on separate
 * streams.  A request's results equal running the requests one by one in admission
 * (= submission) order.  The server owns the tree while it exists: call no other stm_* function
 * on that tree until stm_server_destroy returns. */
typedef struct stm_server stm_server;

typedef struct stm_server_config {
    uint64_t hbm_budget;   /* device bytes for batches in flight; 0 = a quarter of the free HBM,
                              at most 8 GiB; never above 90 % of the free HBM */
    uint32_t slots;        /* batch slots the budget is split into: forming, running, draining
                              (2..8, default 3) */
    uint32_t max_batch_ops;/* ops per batch at most; 0 = a batch takes no request that would
                              carry it past 2^18 ops, and a larger request runs alone (up to
                              what a slot holds) */
    double scan_fraction;  /* share of a slot kept for scan pairs (default 0.5) */
    uint32_t reserved[4];
} stm_server_config;

typedef struct stm_server_stats {
    uint64_t requests, ops, batches, scans, scan_pairs;
    uint64_t max_pending;  /* most requests waiting for admission at once */
    double busy_s;         /* wall time the tree spent running batches */
    uint64_t hbm_budget, slot_ops, slot_pairs;
} stm_server_stats;

int stm_server_config_init(stm_server_config* cfg);
int stm_server_create(stm_tree* tree, const stm_server_config* cfg, stm_server** out);
/* Queues one request; returns at once with a ticket.  The arrays must stay valid (inputs) and
 * untouched (outputs) until stm_server_wait returns for the ticket.  scan_keys / scan_vals hold
 * the request's scan pairs laid out as in stm_run_batch_scan.  arrival_s is the arrival time on
 * the stm_server_now() clock used for the latency (< 0 = now). */
int stm_server_submit(stm_server* srv, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                      size_t n, uint64_t* out_vals, uint8_t* out_found, uint64_t* scan_keys,
                      uint64_t* scan_vals, size_t scan_cap, double arrival_s, uint64_t* ticket);
/* Blocks until the request is done; *latency_s = completion - arrival.  Returns its status. */
int stm_server_wait(stm_server* srv, uint64_t ticket, double* latency_s);
int stm_server_drain(stm_server* srv);
int stm_server_get_stats(stm_server* srv, stm_server_stats* out);
/* Waits for every submitted request, then frees the server (not the tree). */
int stm_server_destroy(stm_server* srv);
double stm_server_now(void);
const char* stm_server_last_error(void);

/* Open-loop load of the serving loop: requests with Poisson arrivals at offered_mops, sizes
 * floor(2^(lo + (hi - lo) u)) ops, lookup_frac lookups of loaded keys, insert_frac inserts of
 * fresh keys, the rest scans of length uniform in [scan_lo, scan_hi].  The requests are generated
 * before the clock starts; a client thread submits each at its arrival time. */
typedef struct stm_serve_workload {
    uint64_t requests;
    double offered_mops;
    uint32_t ops_lo_log2, ops_hi_log2;
    double lookup_frac, insert_frac;
    uint32_t scan_lo, scan_hi;
    uint64_t seed;
    uint32_t warmup_requests; /* requests run first on the same arrival process and left out of
                                 the report's figures (0 = none) */
    uint32_t reserved[3];
} stm_serve_workload;

typedef struct stm_serve_report {
    uint64_t requests, ops, batches, scans, scan_pairs, max_pending, failed;
    double wall_s;          /* first arrival to last completion */
    double mops;            /* ops / wall_s */
    double offered_mops;    /* ops / (last arrival - first arrival) */
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms, mean_ms;  /* request latency */
    double busy_frac;       /* share of the whole run (warm-up included) the tree ran a batch */
    double mean_batch_ops;
    uint64_t hbm_budget, slot_ops;
    uint64_t warmup_requests;  /* requests run before the measured ones */
    double warmup_p99_ms;      /* p99 latency of those */
} stm_serve_report;

int stm_serve_open_loop(stm_tree* tree, const stm_server_config* cfg, const stm_serve_workload* w,
                        const uint64_t* loaded_sorted, size_t n_loaded, stm_serve_report* out,
                        double* latencies_ms);

/* The same open-loop run with every request and answer kept, for checking the answers
 * against a sequential replay in admission order.  On success *record receives the generated
 * requests (warm-up requests included), their answers and their admission tickets; free it
 * with stm_serve_record_free.  No reference counterpart (test and measurement harness). */
typedef struct stm_serve_record {
    uint64_t requests;        /* every request of the run, warm-up ones first */
    uint64_t ops, pairs;      /* ops and scan pairs of all requests */
    uint64_t* req_offsets;    /* [requests + 1]: request r holds ops [req_offsets[r], req_offsets[r+1]) */
    uint64_t* keys;           /* [ops] */
    uint64_t* vals;           /* [ops] insert values; scan lengths for op 3 */
    uint8_t* ops_codes;       /* [ops] */
    uint64_t* out_vals;       /* [ops] answers, as stm_run_batch_scan gives them */
    uint8_t* out_found;       /* [ops] */
    uint64_t* pair_offsets;   /* [requests + 1]: request r's scan pairs */
    uint64_t* scan_keys;      /* [pairs] */
    uint64_t* scan_vals;      /* [pairs] */
    uint64_t* tickets;        /* [requests] admission ticket: requests ran in ticket order */
    void* opaque;
} stm_serve_record;

int stm_serve_open_loop_record(stm_tree* tree, const stm_server_config* cfg,
                               const stm_serve_workload* w, const uint64_t* loaded_sorted,
                               size_t n_loaded, stm_serve_report* out, double* latencies_ms,
                               stm_serve_record** record);
int stm_serve_record_free(stm_serve_record* record);

/* Message of the last failing call on this thread ("" if none). */
const char* stm_last_error(void);

/* Device-buffer plumbing for callers that keep batches resident in HBM (benchmarks,
 * serving loops) without a second GPU runtime in the process.  No reference counterpart:
 * the reference allocates its buffers inside prepareGPU (bpt-gpustm.cpp:196-311). */
int stm_device_alloc(int device, size_t bytes, void** ptr);
int stm_device_free(void* ptr);
int stm_memcpy_h2d(void* dst, const void* src, size_t bytes);
int stm_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned pointers and size); it
 * is also the known-size kernel that HBM traffic counters are scaled by. */
int stm_device_copy(void* dst, const void* src, size_t bytes);
/* Blocks until every kernel queued on the tree's stream has finished. */
int stm_sync(stm_tree* tree);

#ifdef __cplusplus
}
#endif

#endif /* STMTREE_H */
