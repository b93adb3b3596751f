// oracle.cpp — CPU oracle of the STM B+tree batch path.  TEST INFRASTRUCTURE ONLY:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the checker or the reported CPU baseline — never as the product path.
//
// Two independent restatements of the reference's batch semantics:
//
// 1. oracle_run_batch: sequential replay on std::map in input order.  Each op reads the
//    key's current value; INSERT then assigns (src/input-helper.h:27-38: INSERT returns the
//    old value or -1 and updates, GET returns the value or -1).  The reference's v38
//    linearisation (stable sort + in-order per-key fold) is equivalent to this replay.
//    Op codes other than INSERT run as GET, as the reference's merge treats every
//    non-INSERT mission (src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:34).
//
// 2. oracle_pipeline_batch: the reference's own pipeline step by step
//    (src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:289-332): stable sort of (key, position)
//    (:293), run-length encode (:296), exclusive scan of run lengths (:305), mergeMission
//    (bpt-gpustm-kernel-aux.cuh:16-44: per group in order, answer = "forNextAns", which
//    starts as a pointer to the tree answer and becomes the value of each INSERT seen),
//    tree execution of the merged op per unique key (kernel.cu:199-238), writebackAns
//    (aux.cuh:46-56) and the un-permute of test() (bpt-gpustm.cpp:185-192).
//
// Tests require both to agree, and pin them against hand-written known answers,
// brute-force Python dict replays and invariants.
#include <stdint.h>
#include <stddef.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <vector>

namespace {

struct Oracle {
  std::map<uint64_t, uint64_t> m;
};

constexpr uint8_t OP_INSERT = 0;

}  // namespace

extern "C" {

void* oracle_create() { return new Oracle(); }

void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }

uint64_t oracle_size(void* h) { return static_cast<Oracle*>(h)->m.size(); }

// Sequential inserts; the last value of a repeated key wins.
void oracle_bulk_load(void* h, const uint64_t* keys, const uint64_t* vals, size_t n) {
  auto& m = static_cast<Oracle*>(h)->m;
  for (size_t i = 0; i < n; ++i) {
    // A key above every stored key goes to the end in constant time (ascending loads);
    // otherwise a plain assignment (a later duplicate overwrites, as a sequential insert).
    if (m.empty() || m.rbegin()->first < keys[i]) m.emplace_hint(m.end(), keys[i], vals[i]);
    else m[keys[i]] = vals[i];
  }
}

void oracle_run_batch(void* h, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                      size_t n, uint64_t* out_vals, uint8_t* out_found) {
  auto& m = static_cast<Oracle*>(h)->m;
  for (size_t i = 0; i < n; ++i) {
    auto it = m.find(keys[i]);
    const bool found = it != m.end();
    out_found[i] = found ? 1 : 0;
    out_vals[i] = found ? it->second : 0;
    if (ops[i] == OP_INSERT) {
      if (found) it->second = vals[i];
      else m.emplace(keys[i], vals[i]);
    }
  }
}

// Sequential replay with range scans (no reference counterpart: the reference declares RANGE
// but never implements it, src/tree.h:37-39, src/main.cpp:30-34).  Op 3 = SCAN(key, L = vals[i]):
// the first L pairs with key' >= key in the state before op i, written at
// scan_keys/scan_vals[off_i ...] with off_i = sum of the L of the scan ops before i;
// out_vals[i] = pairs written, out_found[i] = (pairs > 0).  Other ops as oracle_run_batch.
// Returns the sum of all L (the scan area the batch needs), or ~0 if it exceeds scan_cap.
uint64_t oracle_run_batch_scan(void* h, const uint64_t* keys, const uint64_t* vals,
                               const uint8_t* ops, size_t n, uint64_t* out_vals,
                               uint8_t* out_found, uint64_t* scan_keys, uint64_t* scan_vals,
                               uint64_t scan_cap) {
  auto& m = static_cast<Oracle*>(h)->m;
  uint64_t off = 0;
  for (size_t i = 0; i < n; ++i) {
    if (ops[i] == 3) {
      const uint64_t L = vals[i];
      if (off + L > scan_cap) return ~0ull;
      uint64_t c = 0;
      for (auto it = m.lower_bound(keys[i]); it != m.end() && c < L; ++it, ++c) {
        scan_keys[off + c] = it->first;
        scan_vals[off + c] = it->second;
      }
      out_vals[i] = c;
      out_found[i] = c > 0 ? 1 : 0;
      off += L;
      continue;
    }
    auto it = m.find(keys[i]);
    const bool found = it != m.end();
    out_found[i] = found ? 1 : 0;
    out_vals[i] = found ? it->second : 0;
    if (ops[i] == OP_INSERT) {
      if (found) it->second = vals[i];
      else m.emplace(keys[i], vals[i]);
    }
  }
  return off;
}

// Restatement of the reference's v38 pipeline on the oracle's map.
void oracle_pipeline_batch(void* h, const uint64_t* keys, const uint64_t* vals, const uint8_t* ops,
                           size_t n, uint64_t* out_vals, uint8_t* out_found) {
  auto& m = static_cast<Oracle*>(h)->m;
  // Phase 1 -> 2: stable sort of (key, original position).
  std::vector<uint32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  // Run-length encode of the sorted keys.
  std::vector<uint64_t> ukeys;
  std::vector<uint32_t> counts;
  for (size_t i = 0; i < n; ++i) {
    const uint64_t k = keys[idx[i]];
    if (ukeys.empty() || ukeys.back() != k) {
      ukeys.push_back(k);
      counts.push_back(0);
    }
    counts.back()++;
  }
  const size_t U = ukeys.size();
  // Exclusive scan of the run lengths.
  std::vector<uint32_t> offs(U, 0);
  for (size_t g = 1; g < U; ++g) offs[g] = offs[g - 1] + counts[g - 1];
  // mergeMission: answers in sorted order; "pending" = the tree's answer of group g.
  struct Ans {
    bool pending;
    uint32_t group;
    uint64_t val;
  };
  std::vector<Ans> sorted_ans(n);
  std::vector<uint8_t> merged_op(U);
  std::vector<uint64_t> merged_val(U);
  for (size_t g = 0; g < U; ++g) {
    Ans next{true, (uint32_t)g, 0};
    uint8_t passed = 1;  // GET
    for (uint32_t i = 0; i < counts[g]; ++i) {
      const uint32_t pos = offs[g] + i;
      const uint32_t orig = idx[pos];
      sorted_ans[pos] = next;
      if (ops[orig] == OP_INSERT) {
        passed = OP_INSERT;
        next = Ans{false, 0, vals[orig]};
      }
    }
    merged_op[g] = passed;
    merged_val[g] = next.val;
  }
  // Tree execution of the merged ops (distinct keys, so their order does not matter).
  std::vector<uint8_t> tree_found(U);
  std::vector<uint64_t> tree_val(U);
  for (size_t g = 0; g < U; ++g) {
    auto it = m.find(ukeys[g]);
    tree_found[g] = it != m.end();
    tree_val[g] = tree_found[g] ? it->second : 0;
    if (merged_op[g] == OP_INSERT) m[ukeys[g]] = merged_val[g];
  }
  // writebackAns + un-permute.
  for (size_t pos = 0; pos < n; ++pos) {
    const uint32_t orig = idx[pos];
    const Ans& a = sorted_ans[pos];
    if (a.pending) {
      out_found[orig] = tree_found[a.group];
      out_vals[orig] = tree_val[a.group];
    } else {
      out_found[orig] = 1;
      out_vals[orig] = a.val;
    }
  }
}

// In-order dump into caller buffers of size oracle_size().
void oracle_dump(void* h, uint64_t* keys, uint64_t* vals) {
  size_t i = 0;
  for (const auto& kv : static_cast<Oracle*>(h)->m) {
    keys[i] = kv.first;
    vals[i] = kv.second;
    ++i;
  }
}

// Restated building blocks, each checked against textbook routines by the tests.
// Stable sort permutation of keys (std::stable_sort).
void oracle_stable_sort_perm(const uint64_t* keys, size_t n, uint32_t* perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
}

}  // extern "C"
