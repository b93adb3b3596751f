// cpu_baseline.cpp — multi-threaded CPU executor of the same batch semantics, timed as the
// reported CPU baseline.  TEST INFRASTRUCTURE ONLY (see oracle.cpp): loaded by tests/ and
// by bench.py's cpu_baseline leg, never by the product path.
//
// The reference has no CPU path.  This executor partitions the key space by T-1
// splitters (quantiles of the bulk-loaded keys); thread t owns partition t as a CPU
// B+tree (64-slot nodes, sorted leaves).  A batch is scattered into per-partition op
// lists that keep input order, then every thread replays its list.  Ops on one key always
// meet in one partition in input order, so results equal the sequential replay of
// oracle.cpp for any T.
#include <stdint.h>
#include <stddef.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

constexpr int NF = 64;  // slots per node

struct Node {
  int n = 0;
  bool leaf = true;
  uint64_t k[NF];
  union {
    uint64_t v[NF];
    Node* c[NF];
  };
};

// B+tree with the min-key convention in inner nodes (k[i] = smallest key under c[i]).
class CpuBTree {
 public:
  ~CpuBTree() { free_all(root_); }

  void bulk_load(const uint64_t* keys, const uint64_t* vals, size_t n, int fill) {
    free_all(root_);
    root_ = nullptr;
    std::vector<Node*> level;
    for (size_t i = 0; i < n || level.empty(); i += fill) {
      Node* l = new Node();
      l->leaf = true;
      const size_t e = std::min(n, i + fill);
      for (size_t j = i; j < e; ++j) {
        l->k[l->n] = keys[j];
        l->v[l->n] = vals[j];
        l->n++;
      }
      level.push_back(l);
      if (n == 0) break;
    }
    while (level.size() > 1) {
      std::vector<Node*> up;
      for (size_t i = 0; i < level.size(); i += fill) {
        Node* p = new Node();
        p->leaf = false;
        const size_t e = std::min(level.size(), i + (size_t)fill);
        for (size_t j = i; j < e; ++j) {
          p->k[p->n] = min_key(level[j]);
          p->c[p->n] = level[j];
          p->n++;
        }
        up.push_back(p);
      }
      level.swap(up);
    }
    root_ = level[0];
  }

  // Appends every (key, value) in key order.
  void dump(std::vector<uint64_t>& ks, std::vector<uint64_t>& vs) const { dump_node(root_, ks, vs); }

  bool get(uint64_t key, uint64_t* val) const {
    const Node* x = root_;
    while (!x->leaf) x = x->c[child_slot(x, key)];
    const int s = lower(x, key);
    if (s < x->n && x->k[s] == key) {
      *val = x->v[s];
      return true;
    }
    return false;
  }

  // Upsert; returns true and the old value if the key was present.
  bool put(uint64_t key, uint64_t val, uint64_t* old) {
    if (root_->n == NF) {
      Node* r = new Node();
      r->leaf = false;
      r->k[0] = min_key(root_);
      r->c[0] = root_;
      r->n = 1;
      split_child(r, 0);
      root_ = r;
    }
    Node* x = root_;
    while (!x->leaf) {
      int s = child_slot(x, key);
      if (x->c[s]->n == NF) {
        split_child(x, s);
        s = child_slot(x, key);
      }
      x = x->c[s];
    }
    const int s = lower(x, key);
    if (s < x->n && x->k[s] == key) {
      *old = x->v[s];
      x->v[s] = val;
      return true;
    }
    std::memmove(x->k + s + 1, x->k + s, (x->n - s) * 8);
    std::memmove(x->v + s + 1, x->v + s, (x->n - s) * 8);
    x->k[s] = key;
    x->v[s] = val;
    x->n++;
    if (s == 0) fix_min(key);
    return false;
  }

 private:
  Node* root_ = nullptr;

  static uint64_t min_key(const Node* x) { return x->n ? x->k[0] : 0; }
  static int lower(const Node* x, uint64_t key) {
    return (int)(std::lower_bound(x->k, x->k + x->n, key) - x->k);
  }
  static int child_slot(const Node* x, uint64_t key) {
    const int s = (int)(std::upper_bound(x->k, x->k + x->n, key) - x->k) - 1;
    return s < 0 ? 0 : s;
  }
  // A new minimum in the leftmost path keeps the min-key separators exact.
  void fix_min(uint64_t key) {
    Node* x = root_;
    while (!x->leaf) {
      if (key < x->k[0]) x->k[0] = key;
      x = x->c[child_slot(x, key)];
    }
  }
  static void split_child(Node* p, int s) {
    Node* a = p->c[s];
    Node* b = new Node();
    b->leaf = a->leaf;
    const int h = a->n / 2;
    b->n = a->n - h;
    std::memcpy(b->k, a->k + h, b->n * 8);
    std::memcpy(b->v, a->v + h, b->n * 8);  // children share the union storage
    a->n = h;
    std::memmove(p->k + s + 2, p->k + s + 1, (p->n - s - 1) * 8);
    std::memmove(p->c + s + 2, p->c + s + 1, (p->n - s - 1) * sizeof(Node*));
    p->k[s + 1] = b->k[0];
    p->c[s + 1] = b;
    p->n++;
  }
  static void dump_node(const Node* x, std::vector<uint64_t>& ks, std::vector<uint64_t>& vs) {
    if (!x) return;
    if (x->leaf) {
      ks.insert(ks.end(), x->k, x->k + x->n);
      vs.insert(vs.end(), x->v, x->v + x->n);
      return;
    }
    for (int i = 0; i < x->n; ++i) dump_node(x->c[i], ks, vs);
  }
  static void free_all(Node* x) {
    if (!x) return;
    if (!x->leaf)
      for (int i = 0; i < x->n; ++i) free_all(x->c[i]);
    delete x;
  }
};

// T persistent worker threads; run(fn) calls fn(t) on every worker t and returns when all
// have finished, so a batch pays no thread start-up.
class Pool {
 public:
  explicit Pool(int T) : T_(T) {
    for (int t = 0; t < T_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& x : th_) x.join();
  }
  void run(const std::function<void(int)>& fn) {
    std::unique_lock<std::mutex> g(m_);
    fn_ = &fn;
    busy_ = T_;
    ++gen_;
    cv_.notify_all();
    done_.wait(g, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void loop(int t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* fn;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        fn = fn_;
      }
      (*fn)(t);
      std::lock_guard<std::mutex> g(m_);
      if (--busy_ == 0) done_.notify_one();
    }
  }
  int T_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* fn_ = nullptr;
  uint64_t gen_ = 0;
  int busy_ = 0;
  bool stop_ = false;
};

struct Executor {
  int T = 1;
  std::vector<uint64_t> splitters;  // T-1 ascending keys
  std::vector<std::unique_ptr<CpuBTree>> parts;
  std::vector<std::vector<uint32_t>> lists;
  std::vector<size_t> cnt, off;
  std::unique_ptr<Pool> pool;

  int part_of(uint64_t key) const {
    return (int)(std::upper_bound(splitters.begin(), splitters.end(), key) - splitters.begin());
  }
};

}  // namespace

extern "C" {

// sorted_keys/vals: the bulk load, ascending unique keys.
void* cpu_exec_create(const uint64_t* sorted_keys, const uint64_t* vals, size_t n, int threads) {
  Executor* e = new Executor();
  e->T = std::max(1, threads);
  const int T = e->T;
  for (int t = 1; t < T; ++t) e->splitters.push_back(n ? sorted_keys[(size_t)n * t / T] : 0);
  e->parts.resize(T);
  e->lists.resize(T);
  e->cnt.assign((size_t)T * T, 0);
  e->off.assign((size_t)T * T, 0);
  e->pool.reset(new Pool(T));
  std::vector<size_t> lo(T + 1);
  for (int t = 0; t <= T; ++t) lo[t] = (size_t)n * t / T;
  e->pool->run([&](int t) {
    e->parts[t].reset(new CpuBTree());
    e->parts[t]->bulk_load(sorted_keys + lo[t], vals + lo[t], lo[t + 1] - lo[t], NF * 3 / 4);
  });
  return e;
}

void cpu_exec_destroy(void* h) { delete static_cast<Executor*>(h); }

// Runs one batch; returns the wall time in seconds of the whole batch (scatter + replay).
double cpu_exec_run_batch(void* h, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                          size_t n, uint64_t* out_vals, uint8_t* out_found) {
  Executor* e = static_cast<Executor*>(h);
  const int T = e->T;
  const auto t0 = std::chrono::steady_clock::now();
  // Per-slice partition counts, then an order-preserving scatter of op indices.
  std::vector<size_t>& cnt = e->cnt;
  std::vector<size_t>& off = e->off;
  std::fill(cnt.begin(), cnt.end(), 0);
  const size_t per = (n + T - 1) / T;
  e->pool->run([&](int s) {
    const size_t b = std::min(n, s * per), en = std::min(n, b + per);
    for (size_t i = b; i < en; ++i) cnt[(size_t)s * T + e->part_of(keys[i])]++;
  });
  for (int p = 0; p < T; ++p) {
    size_t run = 0;
    for (int s = 0; s < T; ++s) {
      off[(size_t)s * T + p] = run;
      run += cnt[(size_t)s * T + p];
    }
    e->lists[p].resize(run);
  }
  e->pool->run([&](int s) {
    const size_t b = std::min(n, s * per), en = std::min(n, b + per);
    size_t* o = &off[(size_t)s * T];
    for (size_t i = b; i < en; ++i) {
      const int p = e->part_of(keys[i]);
      e->lists[p][o[p]++] = (uint32_t)i;
    }
  });
  e->pool->run([&](int p) {
    CpuBTree& tree = *e->parts[p];
    for (uint32_t i : e->lists[p]) {
      uint64_t v = 0;
      bool f;
      if (ops[i] == 0) f = tree.put(keys[i], vals[i], &v);
      else f = tree.get(keys[i], &v);
      out_found[i] = f ? 1 : 0;
      out_vals[i] = f ? v : 0;
    }
  });
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Number of keys held (the sum over the partitions).
uint64_t cpu_exec_size(void* h) {
  Executor* e = static_cast<Executor*>(h);
  std::vector<uint64_t> counts(e->T, 0);
  e->pool->run([&](int p) {
    std::vector<uint64_t> ks, vs;
    e->parts[p]->dump(ks, vs);
    counts[p] = ks.size();
  });
  uint64_t total = 0;
  for (uint64_t c : counts) total += c;
  return total;
}

// In-order dump: the partitions cover ascending key ranges, so their dumps concatenate.
// keys / vals must hold cpu_exec_size() entries.
void cpu_exec_dump(void* h, uint64_t* keys, uint64_t* vals) {
  Executor* e = static_cast<Executor*>(h);
  const int T = e->T;
  std::vector<std::vector<uint64_t>> ks(T), vs(T);
  e->pool->run([&](int p) { e->parts[p]->dump(ks[p], vs[p]); });
  size_t at = 0;
  for (int p = 0; p < T; ++p) {
    std::memcpy(keys + at, ks[p].data(), ks[p].size() * 8);
    std::memcpy(vals + at, vs[p].data(), vs[p].size() * 8);
    at += ks[p].size();
  }
}

}  // extern "C"
