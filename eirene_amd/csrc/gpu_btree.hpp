// gpu_btree.hpp — C++ host interface mirroring the reference's batch API over the C ABI.
//
// Reference: class GPU_BTree (src/tree.h:28-43), Data_Collector (src/input-helper.h:63-218)
// and GPU_STM_Tree (src/gpu-stm-tree-v38/bpt-gpustm.h:166-222).  Same method names and
// meaning: prepareGPU() creates the device tree, prepareInput(file) reads an op file in the
// reference's text format ("R k", "P k v", "G a b"), gpu_work() runs it as batches,
// test() prints every op's answer.  Values of "P" lines go to a host warehouse and the tree
// stores their warehouse offsets, as in the reference (input-helper.h:167-177).  Errors
// throw StmFailure (the reference exits the process, src/cuda_utils.h:3-12).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/stmtree.h"

namespace stm_host {

struct StmFailure : std::runtime_error {
  int code;
  StmFailure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what) {
  if (rc != STM_OK) throw StmFailure(rc, std::string(what) + ": " + stm_last_error());
}

// The reference's keys and values are signed (`int`, src/input-helper.h:70 instantiates
// Data_Collector<int, int>).  Signed keys map to the library's u64 keys by flipping the sign
// bit, which keeps their order (so dumps and ranges come out in signed order); the one key the
// library reserves then corresponds to INT64_MAX, as the reference reserves INT_MAX
// (src/gpu-stm-tree-v38/bpt-gpustm.h:20).
inline uint64_t key_of_signed(int64_t k) { return (uint64_t)k ^ (1ull << 63); }
inline int64_t signed_of_key(uint64_t u) { return (int64_t)(u ^ (1ull << 63)); }

// Longest range (keys) one "G a b" op may span: it runs as consecutive scans of at most
// STM_SCAN_MAX_LEN keys each.
constexpr uint64_t RANGE_MAX_WIDTH = 1ull << 20;

// Op file reader: "R k" -> GET, "P k v" -> INSERT (value stored in the warehouse, the op
// carries its offset), "G a b" -> RANGE over keys a..b (both included).  Lines that do not
// parse are skipped, as in the reference (src/input-helper.h:136-195).
struct DataCollector {
  std::vector<uint64_t> keys;  // mapped keys
  std::vector<uint64_t> vals;  // warehouse offset (INSERT), mapped end key (RANGE), unused (GET)
  std::vector<uint8_t> mission;
  std::vector<int64_t>* warehouse;

  explicit DataCollector(std::vector<int64_t>* wh) : warehouse(wh) {}

  size_t readFile(const std::string& name) {
    std::ifstream f(name);
    if (!f) throw StmFailure(STM_ERR_INVALID, "cannot open " + name);
    std::string line;
    size_t n = 0;
    while (std::getline(f, line)) {
      std::istringstream ss(line);
      char type = 0;
      long long k = 0, v = 0;
      if (!(ss >> type)) continue;
      if (type == 'R') {
        if (!(ss >> k)) continue;
        keys.push_back(key_of_signed(k));
        mission.push_back(STM_OP_GET);
        vals.push_back(0);
      } else if (type == 'P') {
        if (!(ss >> k >> v)) continue;
        keys.push_back(key_of_signed(k));
        mission.push_back(STM_OP_INSERT);
        vals.push_back(warehouse->size());
        warehouse->push_back(v);
      } else if (type == 'G') {
        if (!(ss >> k >> v)) continue;
        keys.push_back(key_of_signed(k));
        mission.push_back(STM_OP_RANGE);
        vals.push_back(key_of_signed(v));
      } else {
        continue;
      }
      ++n;
    }
    return n;
  }

  void clearMissions() {
    keys.clear();
    vals.clear();
    mission.clear();
  }
};

// The pairs a RANGE op found: [first, first + count) of the range pair arrays.
struct RangeAns {
  uint64_t first = 0, count = 0;
};

class GPU_BTree {
 public:
  virtual ~GPU_BTree() {}
  virtual void prepareGPU() = 0;
  virtual void prepareInput(const std::string& fileName) = 0;
  virtual void gpu_work() = 0;
  virtual void test() = 0;
};

// The MI355X STM tree behind the reference's interface.
class GPU_STM_Tree : public GPU_BTree {
 public:
  std::vector<int64_t> warehouse;
  DataCollector dc{&warehouse};
  stm_config cfg;
  stm_tree* tree = nullptr;
  std::vector<uint64_t> ans;
  std::vector<uint8_t> found;
  std::vector<RangeAns> ranges;               // per op (RANGE ops only)
  std::vector<uint64_t> range_keys, range_vals;  // pairs of all RANGE ops, op by op
  FILE* out = stdout;

  explicit GPU_STM_Tree(uint32_t route = STM_ROUTE_DESCENT) {
    stm_config_init(&cfg);
    cfg.route = route;
    cfg.timing = 1;
  }
  ~GPU_STM_Tree() override { stm_destroy(tree); }

  void prepareGPU() override { check(stm_create(&cfg, &tree), "stm_create"); }

  void prepareInput(const std::string& fileName) override {
    dc.clearMissions();
    dc.readFile(fileName);
    ans.assign(dc.keys.size(), 0);
    found.assign(dc.keys.size(), 0);
    ranges.assign(dc.keys.size(), RangeAns{});
    range_keys.clear();
    range_vals.clear();
  }

  void gpu_work() override {
    stm_stats before, after;
    check(stm_get_stats(tree, &before), "stm_get_stats");
    bool any_range = false;
    for (uint8_t m : dc.mission) any_range |= m == STM_OP_RANGE;
    if (any_range) run_with_ranges();
    else
      check(stm_run_batch(tree, dc.keys.data(), dc.vals.data(), dc.mission.data(), dc.keys.size(),
                          ans.data(), found.data()),
            "stm_run_batch");
    check(stm_get_stats(tree, &after), "stm_get_stats");
    double ms = 0;
    for (int p = 0; p < STM_NUM_PHASES; ++p) ms += after.phase_ms[p] - before.phase_ms[p];
    std::fprintf(out, "bucket size\t%llu\n", (unsigned long long)cfg.max_batch);
    std::fprintf(out, "total size\t%zu\n", dc.keys.size());
    std::fprintf(out, "total gpu exec time: %g\n", ms / 1000.0);
    // The reference's ABORT_COUNT totals (src/gpu-stm-tree-v38/bpt-gpustm.cpp:142-160), per
    // file: merged ops ("working num"), aborted attempts, nodes read, leaf-level steps.  The
    // reference keeps them per merged op and sums them on the host; the library sums them on
    // the device.
    const uint64_t working = after.unique_ops - before.unique_ops;
    const uint64_t steps = after.node_reads - before.node_reads;
    std::fprintf(out, "total working num: %llu\n", (unsigned long long)working);
    std::fprintf(out, "total roll back times: %llu\n",
                 (unsigned long long)(after.aborts - before.aborts));
    std::fprintf(out, "total traversal_steps: %llu\n", (unsigned long long)steps);
    std::fprintf(out, "total leaf traversal_steps: %llu\n",
                 (unsigned long long)(after.leaf_steps - before.leaf_steps));
    std::fprintf(out, "average traversal_steps: %g\n", working ? (double)steps / working : 0.0);
  }

  // "G a b" ops: the reference declares RANGE but never runs it (src/tree.h:37-39,
  // src/main.cpp:30-34).  Here each runs as consecutive SCAN ops of at most STM_SCAN_MAX_LEN
  // keys over a..b in the same batch; scans do not change the tree, so every one of them sees
  // the state before the RANGE op, whatever the chunking.  Each scan keeps the pairs of its
  // own key interval, so a pair is reported once.
  void run_with_ranges() {
    std::vector<uint64_t> k, v;
    std::vector<uint8_t> o;
    std::vector<size_t> first_op(dc.keys.size()), n_ops(dc.keys.size());
    for (size_t i = 0; i < dc.keys.size(); ++i) {
      first_op[i] = k.size();
      if (dc.mission[i] != STM_OP_RANGE) {
        k.push_back(dc.keys[i]);
        v.push_back(dc.vals[i]);
        o.push_back(dc.mission[i]);
      } else {
        const uint64_t a = dc.keys[i], b = dc.vals[i];
        if (b >= a && b - a >= RANGE_MAX_WIDTH)
          throw StmFailure(STM_ERR_INVALID, "range op " + std::to_string(i) + " spans more than " +
                                                std::to_string(RANGE_MAX_WIDTH) + " keys");
        for (uint64_t s0 = a; b >= a && s0 <= b && s0 != STM_KEY_RESERVED;) {
          const uint64_t len = std::min<uint64_t>(STM_SCAN_MAX_LEN, b - s0 + 1);
          k.push_back(s0);
          v.push_back(len);
          o.push_back(STM_OP_RANGE);
          if (b - s0 + 1 <= STM_SCAN_MAX_LEN) break;
          s0 += len;
        }
      }
      n_ops[i] = k.size() - first_op[i];
    }
    uint64_t cap = 0;
    std::vector<uint64_t> off(k.size() + 1, 0);
    for (size_t j = 0; j < k.size(); ++j) off[j + 1] = off[j] + (o[j] == STM_OP_RANGE ? v[j] : 0);
    cap = off[k.size()];
    std::vector<uint64_t> ov(k.size()), sk(std::max<uint64_t>(cap, 1)), sv(std::max<uint64_t>(cap, 1));
    std::vector<uint8_t> of(k.size());
    check(stm_run_batch_scan(tree, k.data(), v.data(), o.data(), k.size(), ov.data(), of.data(),
                             sk.data(), sv.data(), cap),
          "stm_run_batch_scan");
    for (size_t i = 0; i < dc.keys.size(); ++i) {
      if (dc.mission[i] != STM_OP_RANGE) {
        ans[i] = ov[first_op[i]];
        found[i] = of[first_op[i]];
        continue;
      }
      RangeAns r;
      r.first = range_keys.size();
      for (size_t j = first_op[i]; j < first_op[i] + n_ops[i]; ++j) {
        // Scan j starts at k[j] and covers keys up to k[j] + v[j] - 1 (at most v[j] keys, so
        // all of them are among its v[j] pairs); the next scan starts after that.
        const uint64_t hi = std::min<uint64_t>(dc.vals[i], k[j] + v[j] - 1);
        for (uint64_t p = 0; p < ov[j]; ++p) {
          if (sk[off[j] + p] > hi) break;
          range_keys.push_back(sk[off[j] + p]);
          range_vals.push_back(sv[off[j] + p]);
        }
      }
      r.count = range_keys.size() - r.first;
      ranges[i] = r;
      ans[i] = r.count;
      found[i] = r.count > 0;
    }
  }

  // Same columns as the reference's Data_Collector::test (src/input-helper.h:206-218): a
  // missing key prints "ans: -1".  A RANGE op prints the number of pairs, then one line per
  // pair ("pair: key\t vals: value") in key order.
  void test() override {
    for (size_t i = 0; i < dc.keys.size(); ++i) {
      const long long key = (long long)signed_of_key(dc.keys[i]);
      if (dc.mission[i] == STM_OP_RANGE) {
        std::fprintf(out, "key: %lld\t operator %u\t ans: %llu\n", key, (unsigned)dc.mission[i],
                     (unsigned long long)ranges[i].count);
        for (uint64_t p = ranges[i].first; p < ranges[i].first + ranges[i].count; ++p)
          std::fprintf(out, "pair: %lld\t vals: %lld\n", (long long)signed_of_key(range_keys[p]),
                       (long long)value_of(range_vals[p]));
      } else if (found[i]) {
        std::fprintf(out, "key: %lld\t operator %u\t ans: %llu\t vals: %lld\n", key,
                     (unsigned)dc.mission[i], (unsigned long long)ans[i], (long long)value_of(ans[i]));
      } else {
        std::fprintf(out, "key: %lld\t operator %u\t ans: -1\n", key, (unsigned)dc.mission[i]);
      }
    }
  }

 private:
  // The tree stores warehouse offsets (src/input-helper.h:167-177); an offset outside the
  // warehouse (a value loaded another way) prints as itself.
  int64_t value_of(uint64_t off) const {
    return off < warehouse.size() ? warehouse[off] : (int64_t)off;
  }
};

}  // namespace stm_host
