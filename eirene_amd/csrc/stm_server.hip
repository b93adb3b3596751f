// stm_server.hip — continuous-batching request loop over one device-resident tree.
//
// The reference runs one op file as a fixed sequence of batches, each copied in, run and
// copied out before the next (GPU_STM_Tree::gpu_work_main,
// src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-135).  Here requests arrive at any time and are
// admitted into the batch being formed as long as its share of the HBM budget holds them:
//
//   slots      the budget is split into `slots` equal batch slots (pinned host staging plus
//              device buffers for keys, values, op codes, answers and scan pairs).  At any
//              time one slot is forming, one is running on the tree, the others drain.
//   admission  FIFO.  A request joins the forming slot if its ops and scan pairs fit;
//              otherwise it and every later request wait until a slot frees up.  Admission
//              reserves the request's place under the server lock; the copy into the slot's
//              pinned buffers happens outside it, in the admitting thread, so several
//              submitting threads copy at once (the runner waits for a slot's copies).
//   runner     when the tree is idle it closes the forming slot (the next free slot starts
//              forming), copies it to HBM on its own stream, runs it as one batch (scan mode
//              when it holds scans) and queues the answer copies on a third stream.
//   completer  waits for a slot's answers, scatters them into the requests' buffers (split
//              over helper threads for large slots), marks the requests done and returns the
//              slot.
// Requests of one batch are concatenated in admission order, and batches run in order, so
// every request sees exactly the state it would see if the requests ran one by one in
// admission order: a result is the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/stmtree.h"

namespace {

thread_local std::string g_server_error;

int sfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_server_error = buf;
  return code;
}

double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

constexpr uint64_t BYTES_PER_OP = 8 + 8 + 1 + 8 + 1;  // key, value, op in; value, found out
constexpr uint64_t BYTES_PER_PAIR = 16;
constexpr unsigned POOL_THREADS = 8;        // helpers for large copy-ins and answer scatters
constexpr uint64_t PARALLEL_BYTES = 4u << 20;  // a copy this large is split over the pool
// Without a max_batch_ops of the caller's, a forming batch takes no request that would carry it
// past this many ops (a larger request still runs alone).  Unbounded batches let a burst grow
// the answer scatter of one batch past what the completer keeps up with: at 200 Mops/s offered
// (2^25 keys, requests of 1..2^16 ops with 2 % scans) unbounded batches gave a p99 of 25-40 ms
// and batches of at most 2^18 ops 7 ms.
constexpr uint64_t SOFT_BATCH_OPS = 1u << 18;

// A fixed set of helper threads that run the parts of one job together with the calling
// thread.  Jobs from different threads run one after another.
class Pool {
 public:
  explicit Pool(unsigned n) {
    for (unsigned i = 0; i < n; ++i) th_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  unsigned size() const { return (unsigned)th_.size() + 1; }
  // Runs f(0) .. f(parts - 1) on the helpers and the caller; returns when all are done.
  void run(unsigned parts, const std::function<void(unsigned)>& f) {
    std::lock_guard<std::mutex> job(job_);
    {
      std::lock_guard<std::mutex> lk(m_);
      f_ = &f;
      parts_ = parts;
      next_ = 0;
      done_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    take();
    std::unique_lock<std::mutex> lk(m_);
    cv_done_.wait(lk, [&] { return done_ == parts_; });
    f_ = nullptr;
  }

 private:
  void take() {
    for (;;) {
      unsigned p;
      const std::function<void(unsigned)>* f;
      {
        std::lock_guard<std::mutex> lk(m_);
        if (!f_ || next_ >= parts_) return;
        p = next_++;
        f = f_;
      }
      (*f)(p);
      bool last;
      {
        std::lock_guard<std::mutex> lk(m_);
        last = ++done_ == parts_;
      }
      if (last) cv_done_.notify_all();
    }
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      take();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_, job_;
  std::condition_variable cv_, cv_done_;
  const std::function<void(unsigned)>* f_ = nullptr;
  unsigned parts_ = 0, next_ = 0, done_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct Req {
  uint64_t ticket = 0;
  const uint64_t* keys = nullptr;
  const uint64_t* vals = nullptr;
  const uint8_t* ops = nullptr;
  size_t n = 0;
  uint64_t* out_vals = nullptr;
  uint8_t* out_found = nullptr;
  uint64_t* scan_keys = nullptr;
  uint64_t* scan_vals = nullptr;
  uint64_t pairs = 0;  // sum of its scan lengths
  uint64_t scans = 0;
  double t_arrival = 0, t_admit = 0, t_done = 0;
  uint64_t off = 0, poff = 0;  // place in its slot (ops, scan pairs)
  int status = STM_OK;
  std::string error;
  bool done = false;
};

struct Slot {
  uint64_t cap_ops = 0, cap_pairs = 0;
  uint64_t *h_keys = nullptr, *h_vals = nullptr, *h_ovals = nullptr, *h_skeys = nullptr,
           *h_svals = nullptr;
  uint8_t *h_ops = nullptr, *h_ofound = nullptr;
  uint64_t *d_keys = nullptr, *d_vals = nullptr, *d_ovals = nullptr, *d_skeys = nullptr,
           *d_svals = nullptr;
  uint8_t *d_ops = nullptr, *d_ofound = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<Req*> reqs;
  uint64_t n = 0, pairs = 0, scans = 0;
  uint32_t copies = 0;  // admitted requests whose inputs are still being copied in
  int status = STM_OK;
  std::string error;

  void reset() {
    reqs.clear();
    n = pairs = scans = 0;
    copies = 0;
    status = STM_OK;
    error.clear();
  }
};

}  // namespace

struct stm_server {
  stm_tree* tree = nullptr;
  int device = 0;
  stm_server_config cfg{};
  uint64_t budget = 0;
  uint64_t soft_ops = SOFT_BATCH_OPS;  // a batch closes to requests that would pass this
  hipStream_t s_in = nullptr, s_out = nullptr, s_tree = nullptr;
  std::vector<std::unique_ptr<Slot>> slots;

  std::mutex mu;
  std::condition_variable cv_run;    // runner: a forming slot has requests, or stop
  std::condition_variable cv_done;   // a request completed
  std::condition_variable cv_drain;  // completer: a slot has answers queued, or stop
  std::deque<Req*> pending;          // waiting for admission
  std::deque<Slot*> free_slots;
  std::deque<Slot*> draining;
  Slot* forming = nullptr;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1;
  uint64_t outstanding = 0;  // submitted, not yet done
  bool stop = false;
  std::thread runner, completer;

  // statistics
  uint64_t st_requests = 0, st_ops = 0, st_batches = 0, st_scans = 0, st_pairs = 0;
  uint64_t st_max_pending = 0;
  double st_busy = 0;

  std::condition_variable cv_copies;  // a slot's input copies finished
  std::unique_ptr<Pool> pool;

  // Admission of queued requests into the forming slot (mu held): reserves their places and
  // returns them in `out`; the caller copies them in (copy_in) after releasing the lock.
  void admit_locked(std::vector<std::pair<Slot*, Req*>>& out) {
    if (!forming && !free_slots.empty()) {
      forming = free_slots.front();
      free_slots.pop_front();
    }
    while (forming && !pending.empty()) {
      Req* r = pending.front();
      Slot* s = forming;
      if (s->n + r->n > s->cap_ops || s->pairs + r->pairs > s->cap_pairs) break;
      if (s->n && s->n + r->n > soft_ops) break;
      pending.pop_front();
      r->off = s->n;
      r->poff = s->pairs;
      s->n += r->n;
      s->pairs += r->pairs;
      s->scans += r->scans;
      s->copies += 1;
      r->t_admit = now_s();
      s->reqs.push_back(r);
      out.emplace_back(s, r);
    }
  }

  // Copies admitted requests into their slots (mu not held), then releases them.
  void copy_in(std::vector<std::pair<Slot*, Req*>>& adm) {
    if (adm.empty()) return;
    uint64_t ops = 0;
    for (auto& sr : adm) ops += sr.second->n;
    // A backlog admitted at once (a slot came free under load) is copied by the pool.
    const unsigned k = ops * 17 > PARALLEL_BYTES && adm.size() > 1 ? pool->size() : 1u;
    auto part = [&adm, k](unsigned q) {
      for (size_t i = q; i < adm.size(); i += k) {
        Slot* s = adm[i].first;
        Req* r = adm[i].second;
        std::memcpy(s->h_keys + r->off, r->keys, r->n * 8);
        std::memcpy(s->h_vals + r->off, r->vals, r->n * 8);
        std::memcpy(s->h_ops + r->off, r->ops, r->n);
      }
    };
    if (k == 1) part(0);
    else pool->run(k, part);
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& sr : adm) sr.first->copies -= 1;
    }
    adm.clear();
    cv_copies.notify_all();
    cv_run.notify_one();
  }

  void run_loop() {
    (void)hipSetDevice(device);
    std::vector<std::pair<Slot*, Req*>> adm;
    for (;;) {
      Slot* s = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_run.wait(lk, [&] { return stop || (forming && forming->n); });
        if (!forming || !forming->n) {
          if (stop) break;
          continue;
        }
        s = forming;
        forming = nullptr;
        admit_locked(adm);  // the next free slot starts forming
        cv_copies.wait(lk, [&] { return s->copies == 0; });  // the slot's inputs are in place
      }
      copy_in(adm);
      const double t0 = now_s();
      hipError_t e = hipMemcpyAsync(s->d_keys, s->h_keys, s->n * 8, hipMemcpyHostToDevice, s_in);
      if (e == hipSuccess) e = hipMemcpyAsync(s->d_vals, s->h_vals, s->n * 8, hipMemcpyHostToDevice, s_in);
      if (e == hipSuccess) e = hipMemcpyAsync(s->d_ops, s->h_ops, s->n, hipMemcpyHostToDevice, s_in);
      if (e == hipSuccess) e = hipEventRecord(s->ev_in, s_in);
      if (e == hipSuccess) e = hipStreamWaitEvent(s_tree, s->ev_in, 0);
      int rc = e == hipSuccess ? STM_OK : sfail(STM_ERR_DEVICE, "request copy: %s", hipGetErrorString(e));
      if (rc == STM_OK) {
        rc = s->scans ? stm_run_batch_scan_device(tree, s->d_keys, s->d_vals, s->d_ops, s->n,
                                                  s->d_ovals, s->d_ofound, s->d_skeys, s->d_svals,
                                                  s->pairs)
                      : stm_run_batch_device(tree, s->d_keys, s->d_vals, s->d_ops, s->n, s->d_ovals,
                                             s->d_ofound);
        if (rc != STM_OK) s->error = stm_last_error();
      } else {
        s->error = g_server_error;
      }
      s->status = rc;
      const double t1 = now_s();
      e = hipMemcpyAsync(s->h_ovals, s->d_ovals, s->n * 8, hipMemcpyDeviceToHost, s_out);
      if (e == hipSuccess) e = hipMemcpyAsync(s->h_ofound, s->d_ofound, s->n, hipMemcpyDeviceToHost, s_out);
      if (e == hipSuccess && s->pairs)
        e = hipMemcpyAsync(s->h_skeys, s->d_skeys, s->pairs * 8, hipMemcpyDeviceToHost, s_out);
      if (e == hipSuccess && s->pairs)
        e = hipMemcpyAsync(s->h_svals, s->d_svals, s->pairs * 8, hipMemcpyDeviceToHost, s_out);
      if (e == hipSuccess) e = hipEventRecord(s->ev_out, s_out);
      if (e != hipSuccess && s->status == STM_OK) {
        s->status = STM_ERR_DEVICE;
        s->error = std::string("answer copy: ") + hipGetErrorString(e);
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        st_batches += 1;
        st_busy += t1 - t0;
        draining.push_back(s);
      }
      cv_drain.notify_one();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      draining.push_back(nullptr);  // end marker for the completer
    }
    cv_drain.notify_one();
  }

  void complete_loop() {
    (void)hipSetDevice(device);
    for (;;) {
      Slot* s = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_drain.wait(lk, [&] { return !draining.empty(); });
        s = draining.front();
        draining.pop_front();
      }
      if (!s) break;
      const hipError_t e = hipEventSynchronize(s->ev_out);
      if (e != hipSuccess && s->status == STM_OK) {
        s->status = STM_ERR_DEVICE;
        s->error = std::string("answer copy: ") + hipGetErrorString(e);
      }
      if (s->status == STM_OK) {
        // Large slots are scattered by the pool (each part takes every k-th request).
        const uint64_t bytes = s->n * 9 + s->pairs * 16;
        const unsigned k = bytes > PARALLEL_BYTES && s->reqs.size() > 1 ? pool->size() : 1u;
        auto scatter = [s, k](unsigned part) {
          for (size_t i = part; i < s->reqs.size(); i += k) {
            Req* r = s->reqs[i];
            std::memcpy(r->out_vals, s->h_ovals + r->off, r->n * 8);
            std::memcpy(r->out_found, s->h_ofound + r->off, r->n);
            if (r->pairs) {
              std::memcpy(r->scan_keys, s->h_skeys + r->poff, r->pairs * 8);
              std::memcpy(r->scan_vals, s->h_svals + r->poff, r->pairs * 8);
            }
          }
        };
        if (k == 1) scatter(0);
        else pool->run(k, scatter);
      }
      const double t = now_s();
      std::vector<std::pair<Slot*, Req*>> adm;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (Req* r : s->reqs) {
          r->status = s->status;
          r->error = s->error;
          r->t_done = t;
          r->done = true;
          st_requests += 1;
          st_ops += r->n;
          st_scans += r->scans;
          st_pairs += r->pairs;
          outstanding -= 1;
        }
        s->reset();
        free_slots.push_back(s);
        admit_locked(adm);
      }
      copy_in(adm);
      cv_done.notify_all();
    }
  }

  void free_all() {
    for (auto& sp : slots) {
      Slot* s = sp.get();
      void* hp[] = {s->h_keys, s->h_vals, s->h_ovals, s->h_skeys, s->h_svals, s->h_ops, s->h_ofound};
      for (void* p : hp)
        if (p) (void)hipHostFree(p);
      void* dp[] = {s->d_keys, s->d_vals, s->d_ovals, s->d_skeys, s->d_svals, s->d_ops, s->d_ofound};
      for (void* p : dp)
        if (p) (void)hipFree(p);
      if (s->ev_in) (void)hipEventDestroy(s->ev_in);
      if (s->ev_out) (void)hipEventDestroy(s->ev_out);
    }
    slots.clear();
    if (s_in) (void)hipStreamDestroy(s_in);
    if (s_out) (void)hipStreamDestroy(s_out);
    s_in = s_out = nullptr;
  }
};

extern "C" {

const char* stm_server_last_error(void) { return g_server_error.c_str(); }

double stm_server_now(void) { return now_s(); }

int stm_server_config_init(stm_server_config* cfg) {
  if (!cfg) return sfail(STM_ERR_INVALID, "null config");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->hbm_budget = 0;
  cfg->slots = 3;
  cfg->max_batch_ops = 0;
  cfg->scan_fraction = 0.5;
  return STM_OK;
}

int stm_server_create(stm_tree* tree, const stm_server_config* cfg_in, stm_server** out) {
  if (!tree || !out) return sfail(STM_ERR_INVALID, "null argument");
  *out = nullptr;
  stm_server_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else stm_server_config_init(&cfg);
  if (cfg.slots < 2 || cfg.slots > 8) return sfail(STM_ERR_INVALID, "slots %u not in [2, 8]", cfg.slots);
  if (!(cfg.scan_fraction > 0 && cfg.scan_fraction < 1))
    return sfail(STM_ERR_INVALID, "scan_fraction must lie in (0, 1)");
  void* stream = nullptr;
  if (stm_get_stream(tree, &stream) != STM_OK) return sfail(STM_ERR_INVALID, "tree stream");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return sfail(STM_ERR_DEVICE, "hipGetDevice");
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return sfail(STM_ERR_DEVICE, "hipMemGetInfo");
  // Admission budget: the caller's, never above 90 % of the HBM free now.
  uint64_t budget = cfg.hbm_budget ? cfg.hbm_budget : std::min<uint64_t>(free_b / 4, 8ull << 30);
  budget = std::min<uint64_t>(budget, (uint64_t)(free_b * 0.9));
  const uint64_t slot_bytes = budget / cfg.slots;
  const uint64_t op_bytes = (uint64_t)(slot_bytes * (1.0 - cfg.scan_fraction));
  uint64_t cap_ops = op_bytes / BYTES_PER_OP;
  if (cfg.max_batch_ops) cap_ops = std::min<uint64_t>(cap_ops, cfg.max_batch_ops);
  const uint64_t cap_pairs = (uint64_t)(slot_bytes * cfg.scan_fraction) / BYTES_PER_PAIR;
  if (cap_ops < 1024 || cap_pairs < STM_SCAN_MAX_LEN)
    return sfail(STM_ERR_NOMEM, "HBM budget of %llu bytes leaves slots too small (%llu ops, %llu pairs)",
                 (unsigned long long)budget, (unsigned long long)cap_ops, (unsigned long long)cap_pairs);
  auto sv = std::make_unique<stm_server>();
  stm_server* s = sv.get();
  s->tree = tree;
  s->device = dev;
  s->cfg = cfg;
  s->budget = budget;
  s->soft_ops = cfg.max_batch_ops ? ~0ull : SOFT_BATCH_OPS;
  s->s_tree = (hipStream_t)stream;
  if (hipStreamCreate(&s->s_in) != hipSuccess || hipStreamCreate(&s->s_out) != hipSuccess) {
    s->free_all();
    return sfail(STM_ERR_DEVICE, "hipStreamCreate");
  }
  for (uint32_t i = 0; i < cfg.slots; ++i) {
    auto sl = std::make_unique<Slot>();
    Slot* p = sl.get();
    s->slots.push_back(std::move(sl));
    p->cap_ops = cap_ops;
    p->cap_pairs = cap_pairs;
    bool ok = hipHostMalloc((void**)&p->h_keys, cap_ops * 8, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_vals, cap_ops * 8, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_ops, cap_ops, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_ovals, cap_ops * 8, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_ofound, cap_ops, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_skeys, cap_pairs * 8, 0) == hipSuccess &&
              hipHostMalloc((void**)&p->h_svals, cap_pairs * 8, 0) == hipSuccess &&
              hipMalloc((void**)&p->d_keys, cap_ops * 8) == hipSuccess &&
              hipMalloc((void**)&p->d_vals, cap_ops * 8) == hipSuccess &&
              hipMalloc((void**)&p->d_ops, cap_ops) == hipSuccess &&
              hipMalloc((void**)&p->d_ovals, cap_ops * 8) == hipSuccess &&
              hipMalloc((void**)&p->d_ofound, cap_ops) == hipSuccess &&
              hipMalloc((void**)&p->d_skeys, cap_pairs * 8) == hipSuccess &&
              hipMalloc((void**)&p->d_svals, cap_pairs * 8) == hipSuccess &&
              hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      s->free_all();
      return sfail(STM_ERR_NOMEM, "server slot %u (%llu ops, %llu scan pairs)", i,
                   (unsigned long long)cap_ops, (unsigned long long)cap_pairs);
    }
    s->free_slots.push_back(p);
  }
  // Copies into and out of pinned host pages that no copy has touched yet ran 5-10 times
  // slower than later ones (measured: the first batches' answer copies took 0.8-1.9 ms against
  // 0.1-0.25 ms later for as many ops), and the first copy in each direction started ≈6 ms late;
  // the first requests queued behind them (a warm-up p99 of 8-16 ms).  So every slot buffer is
  // copied once, whole, in its direction, before any request arrives (≈0.2 s at the default
  // budget of 8 GiB).
  for (auto& sp : s->slots) {
    Slot* p = sp.get();
    if (hipMemcpyAsync(p->d_keys, p->h_keys, cap_ops * 8, hipMemcpyHostToDevice, s->s_in) != hipSuccess ||
        hipMemcpyAsync(p->d_vals, p->h_vals, cap_ops * 8, hipMemcpyHostToDevice, s->s_in) != hipSuccess ||
        hipMemcpyAsync(p->d_ops, p->h_ops, cap_ops, hipMemcpyHostToDevice, s->s_in) != hipSuccess ||
        hipMemcpyAsync(p->h_ovals, p->d_ovals, cap_ops * 8, hipMemcpyDeviceToHost, s->s_out) != hipSuccess ||
        hipMemcpyAsync(p->h_ofound, p->d_ofound, cap_ops, hipMemcpyDeviceToHost, s->s_out) != hipSuccess ||
        hipMemcpyAsync(p->h_skeys, p->d_skeys, cap_pairs * 8, hipMemcpyDeviceToHost, s->s_out) != hipSuccess ||
        hipMemcpyAsync(p->h_svals, p->d_svals, cap_pairs * 8, hipMemcpyDeviceToHost, s->s_out) != hipSuccess) {
      s->free_all();
      return sfail(STM_ERR_DEVICE, "server copy set-up");
    }
  }
  if (hipStreamSynchronize(s->s_in) != hipSuccess || hipStreamSynchronize(s->s_out) != hipSuccess) {
    s->free_all();
    return sfail(STM_ERR_DEVICE, "server copy set-up");
  }
  // One 1-op batch of each kind (a lookup, a scan) through the slot's buffers: the first
  // launch of the scan kernels held the first batch ≈1 ms.  Neither changes the tree.
  {
    Slot* p = s->slots[0].get();
    p->h_keys[0] = 0;
    p->h_vals[0] = 1;
    p->h_ops[0] = STM_OP_RANGE;
    int wrc = hipMemcpy(p->d_keys, p->h_keys, 8, hipMemcpyHostToDevice) == hipSuccess &&
                      hipMemcpy(p->d_vals, p->h_vals, 8, hipMemcpyHostToDevice) == hipSuccess &&
                      hipMemcpy(p->d_ops, p->h_ops, 1, hipMemcpyHostToDevice) == hipSuccess
                  ? STM_OK
                  : STM_ERR_DEVICE;
    if (wrc == STM_OK)
      wrc = stm_run_batch_scan_device(tree, p->d_keys, p->d_vals, p->d_ops, 1, p->d_ovals,
                                      p->d_ofound, p->d_skeys, p->d_svals, 1);
    p->h_ops[0] = STM_OP_GET;
    if (wrc == STM_OK && hipMemcpy(p->d_ops, p->h_ops, 1, hipMemcpyHostToDevice) != hipSuccess)
      wrc = STM_ERR_DEVICE;
    if (wrc == STM_OK)
      wrc = stm_run_batch_device(tree, p->d_keys, p->d_vals, p->d_ops, 1, p->d_ovals, p->d_ofound);
    if (wrc != STM_OK) {
      const std::string why = stm_last_error();
      s->free_all();
      return sfail(wrc, "server warm-up batch: %s", why.c_str());
    }
  }
  s->pool = std::make_unique<Pool>(POOL_THREADS);
  s->runner = std::thread([s] { s->run_loop(); });
  s->completer = std::thread([s] { s->complete_loop(); });
  *out = sv.release();
  return STM_OK;
}

int stm_server_submit(stm_server* s, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                      size_t n, uint64_t* out_vals, uint8_t* out_found, uint64_t* scan_keys,
                      uint64_t* scan_vals, size_t scan_cap, double arrival_s, uint64_t* ticket) {
  if (!s || !ticket) return sfail(STM_ERR_INVALID, "null argument");
  *ticket = 0;
  if (n == 0) return sfail(STM_ERR_INVALID, "empty request");
  if (!keys || !vals || !ops || !out_vals || !out_found) return sfail(STM_ERR_INVALID, "null array");
  uint64_t pairs = 0, scans = 0;
  for (size_t i = 0; i < n; ++i) {
    if (keys[i] == STM_KEY_RESERVED)
      return sfail(STM_ERR_INVALID, "op %zu uses the reserved key 0xffffffffffffffff", i);
    if (ops[i] > STM_OP_RANGE) return sfail(STM_ERR_INVALID, "op %zu has op code %u", i, ops[i]);
    if (ops[i] == STM_OP_RANGE) {
      if (vals[i] < 1 || vals[i] > STM_SCAN_MAX_LEN)
        return sfail(STM_ERR_INVALID, "scan op %zu has length %llu, not in [1, %u]", i,
                     (unsigned long long)vals[i], STM_SCAN_MAX_LEN);
      pairs += vals[i];
      scans += 1;
    }
  }
  if (pairs > scan_cap || (pairs && (!scan_keys || !scan_vals)))
    return sfail(STM_ERR_INVALID, "request scans need %llu pairs, its area holds %zu",
                 (unsigned long long)pairs, scan_cap);
  if (n > s->slots[0]->cap_ops || pairs > s->slots[0]->cap_pairs)
    return sfail(STM_ERR_NOMEM, "request of %zu ops and %llu scan pairs exceeds a batch slot "
                 "(%llu ops, %llu pairs); raise hbm_budget", n, (unsigned long long)pairs,
                 (unsigned long long)s->slots[0]->cap_ops, (unsigned long long)s->slots[0]->cap_pairs);
  auto r = std::make_unique<Req>();
  r->keys = keys;
  r->vals = vals;
  r->ops = ops;
  r->n = n;
  r->out_vals = out_vals;
  r->out_found = out_found;
  r->scan_keys = scan_keys;
  r->scan_vals = scan_vals;
  r->pairs = pairs;
  r->scans = scans;
  r->t_arrival = arrival_s >= 0 ? arrival_s : now_s();
  std::vector<std::pair<Slot*, Req*>> adm;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return sfail(STM_ERR_STATE, "server is shutting down");
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    Req* rp = r.get();
    s->reqs.emplace(rp->ticket, std::move(r));
    s->pending.push_back(rp);
    s->outstanding += 1;
    s->st_max_pending = std::max<uint64_t>(s->st_max_pending, s->pending.size());
    s->admit_locked(adm);
  }
  s->copy_in(adm);
  return STM_OK;
}

int stm_server_wait(stm_server* s, uint64_t ticket, double* latency_s) {
  if (!s) return sfail(STM_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return sfail(STM_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  Req* r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->done; });
  if (latency_s) *latency_s = r->t_done - r->t_arrival;
  const int rc = r->status;
  if (rc != STM_OK) g_server_error = r->error;
  s->reqs.erase(it);
  return rc;
}

int stm_server_drain(stm_server* s) {
  if (!s) return sfail(STM_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
  return STM_OK;
}

int stm_server_get_stats(stm_server* s, stm_server_stats* out) {
  if (!s || !out) return sfail(STM_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->requests = s->st_requests;
  out->ops = s->st_ops;
  out->batches = s->st_batches;
  out->scans = s->st_scans;
  out->scan_pairs = s->st_pairs;
  out->max_pending = s->st_max_pending;
  out->busy_s = s->st_busy;
  out->hbm_budget = s->budget;
  out->slot_ops = s->slots.empty() ? 0 : s->slots[0]->cap_ops;
  out->slot_pairs = s->slots.empty() ? 0 : s->slots[0]->cap_pairs;
  return STM_OK;
}

int stm_server_destroy(stm_server* s) {
  if (!s) return STM_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    s->stop = true;
  }
  s->cv_run.notify_all();
  if (s->runner.joinable()) s->runner.join();
  if (s->completer.joinable()) s->completer.join();
  (void)hipSetDevice(s->device);
  s->free_all();
  delete s;
  return STM_OK;
}

// Workload generators of stm_gen.cpp (exported by the same library).
int stm_gen_serving_request(const uint64_t* loaded_sorted, size_t n_loaded, uint64_t seed,
                            uint64_t req, size_t n_ops, double lookup_frac, double insert_frac,
                            uint32_t scan_lo, uint32_t scan_hi, uint64_t* keys, uint64_t* vals,
                            uint8_t* ops);
void stm_gen_batch_sizes(size_t count, uint64_t seed, double lo_exp, double hi_exp, uint64_t* out);
uint64_t stm_splitmix64(uint64_t x);

}  // extern "C"

namespace {

// What a recorded open-loop run keeps: the generated requests, their answers and their
// admission tickets.  The C view points into the vectors.
struct RecordHolder {
  stm_serve_record c{};
  std::vector<uint64_t> off, keys, vals, ovals, poff, skeys, svals, tickets;
  std::vector<uint8_t> ops, ofound;
};

int open_loop_impl(stm_tree* tree, const stm_server_config* cfg, const stm_serve_workload* w,
                   const uint64_t* loaded_sorted, size_t n_loaded, stm_serve_report* out,
                   double* latencies_ms, stm_serve_record** record) {
  if (record) *record = nullptr;
  if (!tree || !w || !out || !loaded_sorted || !n_loaded) return sfail(STM_ERR_INVALID, "null argument");
  // The first W requests warm the loop up (device clocks, host threads) on the same arrival
  // process; the latency figures cover the R requests after them.
  const size_t W = (size_t)w->warmup_requests;
  const size_t R = (size_t)w->requests;
  if (R == 0 || !(w->offered_mops > 0)) return sfail(STM_ERR_INVALID, "no load to offer");
  const size_t RT = R + W;
  std::memset(out, 0, sizeof(*out));
  // Request sizes, arrival times (Poisson at offered_mops), then the request data.
  std::vector<uint64_t> sizes(RT);
  stm_gen_batch_sizes(RT, w->seed, w->ops_lo_log2, w->ops_hi_log2, sizes.data());
  std::vector<uint64_t> off(RT + 1, 0);
  for (size_t r = 0; r < RT; ++r) off[r + 1] = off[r] + std::max<uint64_t>(sizes[r], 1);
  const double mean_ops = (double)off[RT] / (double)RT;
  const double lambda = w->offered_mops * 1e6 / mean_ops;  // requests per second
  std::vector<double> arrival(RT);
  double t = 0;
  for (size_t r = 0; r < RT; ++r) {
    const uint64_t h = stm_splitmix64(w->seed ^ stm_splitmix64(0xa441 + r));
    const double u = ((h >> 11) + 1) * (1.0 / 9007199254740992.0);
    t += -std::log(u) / lambda;
    arrival[r] = t;
  }
  const uint64_t N = off[RT];
  std::vector<uint64_t> keys(N), vals(N), ovals(N);
  std::vector<uint8_t> ops(N), ofound(N);
  {
    const unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::atomic<int> bad{0};
    for (unsigned k = 0; k < T; ++k)
      th.emplace_back([&, k] {
        for (size_t r = k; r < RT; r += T)
          if (stm_gen_serving_request(loaded_sorted, n_loaded, w->seed, r, off[r + 1] - off[r],
                                      w->lookup_frac, w->insert_frac, w->scan_lo, w->scan_hi,
                                      keys.data() + off[r], vals.data() + off[r], ops.data() + off[r]))
            bad = 1;
      });
    for (auto& x : th) x.join();
    if (bad) return sfail(STM_ERR_INVALID, "serving workload parameters");
  }
  std::vector<uint64_t> poff(RT + 1, 0);
  for (size_t r = 0; r < RT; ++r) {
    uint64_t p = 0;
    for (uint64_t i = off[r]; i < off[r + 1]; ++i)
      if (ops[i] == STM_OP_RANGE) p += vals[i];
    poff[r + 1] = poff[r] + p;
  }
  std::vector<uint64_t> skeys(std::max<uint64_t>(poff[RT], 1)), svals(std::max<uint64_t>(poff[RT], 1));

  stm_server* srv = nullptr;
  int rc = stm_server_create(tree, cfg, &srv);
  if (rc) return rc;
  // CLIENTS threads submit the requests (client c takes every CLIENTS-th), each at its
  // arrival time, so copying requests in does not delay later arrivals.
  constexpr unsigned CLIENTS = 4;
  std::vector<uint64_t> tickets(RT, 0);
  std::atomic<int> submit_rc{STM_OK};
  const double t0 = now_s() + 0.005;
  {
    std::vector<std::thread> clients;
    for (unsigned cl = 0; cl < CLIENTS; ++cl)
      clients.emplace_back([&, cl] {
        for (size_t r = cl; r < RT && submit_rc.load() == STM_OK; r += CLIENTS) {
          const double due = t0 + arrival[r];
          for (double n = now_s(); n < due; n = now_s())
            if (due - n > 300e-6) std::this_thread::sleep_for(std::chrono::microseconds(100));
          const int rcs = stm_server_submit(srv, keys.data() + off[r], vals.data() + off[r],
                                            ops.data() + off[r], off[r + 1] - off[r],
                                            ovals.data() + off[r], ofound.data() + off[r],
                                            skeys.data() + poff[r], svals.data() + poff[r],
                                            poff[r + 1] - poff[r], due, &tickets[r]);
          if (rcs != STM_OK) submit_rc = rcs;
        }
      });
    for (auto& x : clients) x.join();
  }
  rc = submit_rc.load();
  std::vector<double> lat, warm;
  lat.reserve(R);
  warm.reserve(W);
  uint64_t failed = 0;
  for (size_t r = 0; r < RT; ++r) {
    if (!tickets[r]) continue;
    double l = 0;
    if (stm_server_wait(srv, tickets[r], &l) != STM_OK) failed += 1;
    if (r < W) {
      warm.push_back(l * 1e3);
      continue;
    }
    lat.push_back(l * 1e3);
    if (latencies_ms) latencies_ms[r - W] = l * 1e3;
  }
  const double t_end = now_s();
  stm_server_stats st{};
  stm_server_get_stats(srv, &st);
  stm_server_destroy(srv);
  if (rc) return rc;
  std::vector<double> sorted_lat = lat;
  std::sort(sorted_lat.begin(), sorted_lat.end());
  auto pct_of = [](const std::vector<double>& v, double q) {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, (size_t)std::floor(q * (v.size() - 1) + 0.5));
    return v[i];
  };
  auto pct = [&](double q) { return pct_of(sorted_lat, q); };
  std::sort(warm.begin(), warm.end());
  double sum = 0;
  for (double x : lat) sum += x;
  // Measured window: from the first measured arrival to the last completion.
  const double t_first = t0 + arrival[W];
  out->requests = R;
  out->ops = N - off[W];
  out->batches = st.batches;
  out->scans = st.scans;
  out->scan_pairs = st.scan_pairs;
  out->max_pending = st.max_pending;
  out->failed = failed;
  out->wall_s = t_end - t_first;
  out->mops = out->ops / out->wall_s / 1e6;
  out->offered_mops = R > 1 ? out->ops / (arrival[RT - 1] - arrival[W]) / 1e6 : 0;
  out->p50_ms = pct(0.50);
  out->p90_ms = pct(0.90);
  out->p99_ms = pct(0.99);
  out->p999_ms = pct(0.999);
  out->max_ms = sorted_lat.empty() ? 0 : sorted_lat.back();
  out->mean_ms = lat.empty() ? 0 : sum / lat.size();
  out->busy_frac = st.busy_s / (t_end - t0);  // over the whole run, warm-up included
  out->warmup_requests = W;
  out->warmup_p99_ms = pct_of(warm, 0.99);
  out->mean_batch_ops = st.batches ? (double)st.ops / st.batches : 0;
  out->hbm_budget = st.hbm_budget;
  out->slot_ops = st.slot_ops;
  if (failed) return sfail(STM_ERR_DEVICE, "%llu requests failed", (unsigned long long)failed);
  if (record) {
    auto h = std::make_unique<RecordHolder>();
    h->off = std::move(off);
    h->keys = std::move(keys);
    h->vals = std::move(vals);
    h->ops = std::move(ops);
    h->ovals = std::move(ovals);
    h->ofound = std::move(ofound);
    h->poff = std::move(poff);
    h->skeys = std::move(skeys);
    h->svals = std::move(svals);
    h->tickets = std::move(tickets);
    stm_serve_record& c = h->c;
    c.requests = RT;
    c.ops = N;
    c.pairs = h->poff[RT];
    c.req_offsets = h->off.data();
    c.keys = h->keys.data();
    c.vals = h->vals.data();
    c.ops_codes = h->ops.data();
    c.out_vals = h->ovals.data();
    c.out_found = h->ofound.data();
    c.pair_offsets = h->poff.data();
    c.scan_keys = h->skeys.data();
    c.scan_vals = h->svals.data();
    c.tickets = h->tickets.data();
    c.opaque = h.get();
    *record = &h.release()->c;
  }
  return STM_OK;
}

}  // namespace

extern "C" {

int stm_serve_open_loop(stm_tree* tree, const stm_server_config* cfg, const stm_serve_workload* w,
                        const uint64_t* loaded_sorted, size_t n_loaded, stm_serve_report* out,
                        double* latencies_ms) {
  return open_loop_impl(tree, cfg, w, loaded_sorted, n_loaded, out, latencies_ms, nullptr);
}

int stm_serve_open_loop_record(stm_tree* tree, const stm_server_config* cfg,
                               const stm_serve_workload* w, const uint64_t* loaded_sorted,
                               size_t n_loaded, stm_serve_report* out, double* latencies_ms,
                               stm_serve_record** record) {
  if (!record) return sfail(STM_ERR_INVALID, "null record");
  return open_loop_impl(tree, cfg, w, loaded_sorted, n_loaded, out, latencies_ms, record);
}

int stm_serve_record_free(stm_serve_record* record) {
  if (record) delete static_cast<RecordHolder*>(record->opaque);
  return STM_OK;
}

}  // extern "C"
