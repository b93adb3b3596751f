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

// Op file reader: "R k" -> GET, "P k v" -> INSERT (value stored in the warehouse, the op
// carries its offset), "G a b" -> RANGE (start a, end b).  Lines that do not parse are
// skipped, as in the reference.
struct DataCollector {
  std::vector<uint64_t> keys;
  std::vector<uint64_t> vals;  // warehouse offset (INSERT), end key (RANGE), unused (GET)
  std::vector<uint8_t> mission;
  std::vector<uint64_t>* warehouse;

  explicit DataCollector(std::vector<uint64_t>* wh) : warehouse(wh) {}

  size_t readFile(const std::string& name) {
    std::ifstream f(name);
    if (!f) throw StmFailure(STM_ERR_INVALID, "cannot open " + name);
    std::string line;
    size_t n = 0;
    while (std::getline(f, line)) {
      std::istringstream ss(line);
      char type = 0;
      uint64_t k = 0, v = 0;
      if (!(ss >> type)) continue;
      if (type == 'R') {
        if (!(ss >> k)) continue;
        keys.push_back(k);
        mission.push_back(STM_OP_GET);
        vals.push_back(0);
      } else if (type == 'P') {
        if (!(ss >> k >> v)) continue;
        keys.push_back(k);
        mission.push_back(STM_OP_INSERT);
        vals.push_back(warehouse->size());
        warehouse->push_back(v);
      } else if (type == 'G') {
        if (!(ss >> k >> v)) continue;
        keys.push_back(k);
        mission.push_back(STM_OP_RANGE);
        vals.push_back(v);
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

  // Same columns as the reference's Data_Collector::test (input-helper.h:206-218); a missing
  // key prints "ans: -1".
  void test(const std::vector<uint64_t>& ans, const std::vector<uint8_t>& found, FILE* out) const {
    for (size_t i = 0; i < keys.size(); ++i) {
      if (found[i])
        std::fprintf(out, "key: %llu\t operator %u\t ans: %llu\t vals: %llu\n",
                     (unsigned long long)keys[i], (unsigned)mission[i], (unsigned long long)ans[i],
                     (unsigned long long)(ans[i] < warehouse->size() ? (*warehouse)[ans[i]] : ans[i]));
      else
        std::fprintf(out, "key: %llu\t operator %u\t ans: -1\n", (unsigned long long)keys[i],
                     (unsigned)mission[i]);
    }
  }
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
  std::vector<uint64_t> warehouse;
  DataCollector dc{&warehouse};
  stm_config cfg;
  stm_tree* tree = nullptr;
  std::vector<uint64_t> ans;
  std::vector<uint8_t> found;
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
  }

  void gpu_work() override {
    stm_stats before, after;
    check(stm_get_stats(tree, &before), "stm_get_stats");
    check(stm_run_batch(tree, dc.keys.data(), dc.vals.data(), dc.mission.data(), dc.keys.size(),
                        ans.data(), found.data()),
          "stm_run_batch");
    check(stm_get_stats(tree, &after), "stm_get_stats");
    double ms = 0;
    for (int p = 0; p < STM_NUM_PHASES; ++p) ms += after.phase_ms[p] - before.phase_ms[p];
    std::fprintf(out, "bucket size\t%llu\n", (unsigned long long)cfg.max_batch);
    std::fprintf(out, "total size\t%zu\n", dc.keys.size());
    std::fprintf(out, "total gpu exec time: %g\n", ms / 1000.0);
    // The reference's ABORT_COUNT totals (src/gpu-stm-tree-v38/bpt-gpustm.cpp:142-160), per batch.
    std::fprintf(out, "total roll back times: %llu\n",
                 (unsigned long long)(after.aborts - before.aborts));
    std::fprintf(out, "total traversal_steps: %llu\n",
                 (unsigned long long)(after.node_reads - before.node_reads));
  }

  void test() override { dc.test(ans, found, out); }
};

}  // namespace stm_host
