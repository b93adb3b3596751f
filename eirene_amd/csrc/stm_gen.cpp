// stm_gen.cpp — seeded synthetic workloads for the batch path (host C++).
//
// All randomness is counter based: draw i of stream s under seed S is
// splitmix64(S ^ (s * GOLDEN) + i * GOLDEN2), so results do not depend on the number of
// threads.  Bulk-load values are splitmix64(key); the values a batch inserts are
// splitmix64(key ^ salt(seed)), so an insert of a key that is already stored carries a value
// different from the stored one and an answer that returns the new value instead of the old
// one is visible.  Distributions:
//   uniform  keys in [0, 2^63)
//   normal   N(p0, p1) by Box-Muller, clamped to [0, 2^63)
//   gamma    Gamma(k = p0 (integer), theta = p1) as a sum of k exponentials, clamped
// The reference has no generator (its op files are absent, .gitignore:49); its text
// format is "R k" / "P k v" / "G a b" (src/input-helper.h:136-195).
#include <stdint.h>
#include <stddef.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t GOLDEN = 0x9e3779b97f4a7c15ull;
constexpr uint64_t KEY_LIMIT = 1ull << 63;

inline uint64_t splitmix64(uint64_t x) {
  x += GOLDEN;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
  return splitmix64(seed ^ splitmix64(stream * 0xd1b54a32d192ed03ull) ^ (i * 0xa0761d6478bd642full));
}

inline double unit(uint64_t r) {  // (0, 1]
  return ((r >> 11) + 1) * (1.0 / 9007199254740992.0);
}

int hw_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 4;
  return (int)std::min(t, 16u);
}

template <typename Fn>
void parallel_for(size_t n, Fn fn) {
  const int T = hw_threads();
  if (n < 65536 || T == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + T - 1) / T;
  for (int t = 0; t < T; ++t) {
    const size_t b = t * per, e = std::min(n, b + per);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

// Parallel LSD radix sort of u64 (8 passes of 8 bits).
void psort_u64(std::vector<uint64_t>& v) {
  const size_t n = v.size();
  if (n < (1u << 16)) {
    std::sort(v.begin(), v.end());
    return;
  }
  const int T = hw_threads();
  std::vector<uint64_t> tmp(n);
  uint64_t* src = v.data();
  uint64_t* dst = tmp.data();
  const size_t per = (n + T - 1) / T;
  std::vector<size_t> cnt((size_t)T * 256);
  for (int pass = 0; pass < 8; ++pass) {
    const int sh = pass * 8;
    std::fill(cnt.begin(), cnt.end(), 0);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] {
        const size_t b = t * per, e = std::min(n, b + per);
        size_t* c = &cnt[(size_t)t * 256];
        for (size_t i = b; i < e; ++i) c[(src[i] >> sh) & 255]++;
      });
    for (auto& x : th) x.join();
    size_t run = 0;
    for (int d = 0; d < 256; ++d)
      for (int t = 0; t < T; ++t) {
        size_t c = cnt[(size_t)t * 256 + d];
        cnt[(size_t)t * 256 + d] = run;
        run += c;
      }
    th.clear();
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] {
        const size_t b = t * per, e = std::min(n, b + per);
        size_t* c = &cnt[(size_t)t * 256];
        for (size_t i = b; i < e; ++i) dst[c[(src[i] >> sh) & 255]++] = src[i];
      });
    for (auto& x : th) x.join();
    std::swap(src, dst);
  }
  // 8 passes: data is back in v.
}

uint64_t sample_key(int dist, double p0, double p1, uint64_t seed, uint64_t stream, uint64_t i) {
  switch (dist) {
    case 1: {  // normal
      const double u1 = unit(draw(seed, stream, 2 * i));
      const double u2 = unit(draw(seed, stream, 2 * i + 1));
      const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      double x = p0 + p1 * z;
      if (!(x > 0)) x = 0;
      if (x >= 9.2233720368547748e18) return KEY_LIMIT - 1;
      return (uint64_t)x;
    }
    case 2: {  // gamma, integer shape
      const int k = std::max(1, (int)std::lround(p0));
      double s = 0;
      for (int j = 0; j < k; ++j) s -= std::log(unit(draw(seed, stream, (uint64_t)k * i + j)));
      double x = p1 * s;
      if (x >= 9.2233720368547748e18) return KEY_LIMIT - 1;
      return (uint64_t)x;
    }
    default:
      return draw(seed, stream, i) >> 1;
  }
}

bool contains(const uint64_t* sorted, size_t n, uint64_t k) {
  return std::binary_search(sorted, sorted + n, k);
}

}  // namespace

extern "C" {

uint64_t stm_splitmix64(uint64_t x) { return splitmix64(x); }

// Insert values of a generated batch with seed `seed`: splitmix64(key ^ salt), salt != 0.
uint64_t stm_batch_value_salt(uint64_t seed) { return draw(seed, 13, 0) | 1; }

void stm_gen_batch_values(const uint64_t* keys, size_t n, uint64_t seed, uint64_t* out) {
  const uint64_t salt = stm_batch_value_salt(seed);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) out[i] = splitmix64(keys[i] ^ salt);
  });
}

void stm_gen_values(const uint64_t* keys, size_t n, uint64_t* out) {
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) out[i] = splitmix64(keys[i]);
  });
}

// n unique keys of the distribution, in a seeded random order.  Returns 0 on success.
int stm_gen_keys(int dist, size_t n, uint64_t seed, double p0, double p1, uint64_t* out) {
  std::vector<uint64_t> v;
  uint64_t drawn = 0;
  size_t want = n + n / 16 + 64;
  for (int round = 0; round < 64; ++round) {
    const size_t old = v.size();
    v.resize(old + want);
    const uint64_t base = drawn;
    parallel_for(want, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) v[old + i] = sample_key(dist, p0, p1, seed, 1, base + i);
    });
    drawn += want;
    psort_u64(v);
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (!v.empty() && v.back() == ~0ull) v.pop_back();
    if (v.size() >= n) break;
    want = (n - v.size()) * 2 + 64;
  }
  if (v.size() < n) return 1;
  // Keep n keys chosen by a seeded shuffle of the unique set, then shuffle their order.
  std::vector<uint64_t> order(v.size());
  for (size_t i = 0; i < v.size(); ++i) order[i] = i;
  for (size_t i = v.size(); i > 1; --i) {
    const size_t j = draw(seed, 2, i) % i;
    std::swap(order[i - 1], order[j]);
  }
  for (size_t i = 0; i < n; ++i) out[i] = v[order[i]];
  return 0;
}

// Mixed batch over a loaded key set (sorted ascending, unique).
//   lookups  = floor(n_ops * lookup_frac); hits = round(lookups * hit_frac) sampled uniformly
//   (with replacement) from the loaded keys, the other lookups are fresh keys of the
//   distribution not in the loaded set; the rest of the ops are inserts of fresh keys.
// Op order is a seeded Fisher-Yates shuffle.  vals = splitmix64(key ^ salt(seed)).
int stm_gen_mixed_batch(const uint64_t* loaded_sorted, size_t n_loaded, size_t n_ops, uint64_t seed,
                        double lookup_frac, double hit_frac, int dist, double p0, double p1,
                        uint64_t* keys, uint64_t* vals, uint8_t* ops) {
  if (n_loaded == 0 && hit_frac > 0) return 1;
  const size_t n_look = (size_t)std::floor((double)n_ops * lookup_frac);
  const size_t n_hit = (size_t)std::llround((double)n_look * hit_frac);
  parallel_for(n_ops, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      uint64_t k;
      if (i < n_hit) {
        k = loaded_sorted[draw(seed, 10, i) % n_loaded];
      } else {
        uint64_t j = 0;
        do {
          k = sample_key(dist, p0, p1, seed, 11 + (j << 8), i);
          ++j;
        } while ((contains(loaded_sorted, n_loaded, k) || k == ~0ull) && j < 1000);
      }
      keys[i] = k;
      ops[i] = i < n_look ? 1 : 0;
    }
  });
  for (size_t i = n_ops; i > 1; --i) {
    const size_t j = draw(seed, 12, i) % i;
    std::swap(keys[i - 1], keys[j]);
    std::swap(ops[i - 1], ops[j]);
  }
  stm_gen_batch_values(keys, n_ops, seed, vals);
  return 0;
}

// Zipf(theta) ranks over [0, n) by the method of Gray et al. (SIGMOD 1994, "Quickly
// generating billion-record synthetic databases"); rank 0 is the most frequent.
struct ZipfGen {
  double theta, zetan, alpha, eta, half_pow;
  uint64_t n;
  ZipfGen(uint64_t n_, double th) : theta(th), n(n_) {
    zetan = 0;
    for (uint64_t i = 1; i <= n; ++i) zetan += 1.0 / std::pow((double)i, theta);
    const double zeta2 = 1.0 + std::pow(0.5, theta);
    alpha = 1.0 / (1.0 - theta);
    eta = (1.0 - std::pow(2.0 / (double)n, 1.0 - theta)) / (1.0 - zeta2 / zetan);
    half_pow = std::pow(0.5, theta);
  }
  uint64_t operator()(double u) const {
    const double uz = u * zetan;
    if (uz < 1.0) return 0;
    if (uz < 1.0 + half_pow) return 1;
    uint64_t r = (uint64_t)((double)n * std::pow(eta * u - eta + 1.0, alpha));
    return r >= n ? n - 1 : r;
  }
};

// Contended batch: ranks ~ Zipf(theta) over the sorted loaded keys; lookups hit
// sorted[rank]; inserts use sorted[rank] + 1 + (h mod (gap - 1)) with gap the distance to
// the next loaded key, so they land in the hot leaves (an update when gap < 2).
int stm_gen_zipf_batch(const uint64_t* loaded_sorted, size_t n_loaded, size_t n_ops, uint64_t seed,
                       double theta, double lookup_frac, uint64_t* keys, uint64_t* vals, uint8_t* ops) {
  if (n_loaded < 2) return 1;
  static thread_local uint64_t cached_n = 0;
  static thread_local double cached_theta = 0;
  static thread_local ZipfGen* gen = nullptr;
  if (!gen || cached_n != n_loaded || cached_theta != theta) {
    delete gen;
    gen = new ZipfGen(n_loaded, theta);
    cached_n = n_loaded;
    cached_theta = theta;
  }
  const ZipfGen& z = *gen;
  parallel_for(n_ops, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const uint64_t r = z(unit(draw(seed, 20, i)));
      const bool look = unit(draw(seed, 21, i)) <= lookup_frac;
      uint64_t k = loaded_sorted[r];
      if (!look) {
        const uint64_t nxt = r + 1 < n_loaded ? loaded_sorted[r + 1] : KEY_LIMIT;
        const uint64_t gap = nxt - k;
        if (gap >= 2) k = k + 1 + draw(seed, 22, i) % (gap - 1);
      }
      keys[i] = k;
      ops[i] = look ? 1 : 0;
    }
  });
  stm_gen_batch_values(keys, n_ops, seed, vals);
  return 0;
}

// One serving-loop request: n_ops ops, lookup_frac lookups (hits on loaded keys), insert_frac
// inserts of fresh uniform keys, the rest range scans from a uniform start key with length
// uniform in [scan_lo, scan_hi] (held in vals).  Request r of seed S draws from its own
// streams, so requests can be generated in any order.
int stm_gen_serving_request(const uint64_t* loaded_sorted, size_t n_loaded, uint64_t seed,
                            uint64_t req, size_t n_ops, double lookup_frac, double insert_frac,
                            uint32_t scan_lo, uint32_t scan_hi, uint64_t* keys, uint64_t* vals,
                            uint8_t* ops) {
  if (n_loaded == 0 || scan_lo == 0 || scan_hi < scan_lo) return 1;
  const uint64_t rs = splitmix64(seed ^ (req * GOLDEN));
  const uint64_t salt = stm_batch_value_salt(rs);
  for (size_t i = 0; i < n_ops; ++i) {
    const double u = unit(draw(rs, 40, i));
    const uint64_t r = draw(rs, 41, i);
    if (u <= lookup_frac) {
      keys[i] = loaded_sorted[r % n_loaded];
      vals[i] = 0;
      ops[i] = 1;
    } else if (u <= lookup_frac + insert_frac) {
      keys[i] = r & (KEY_LIMIT - 1);
      vals[i] = splitmix64(keys[i] ^ salt);
      ops[i] = 0;
    } else {
      keys[i] = r & (KEY_LIMIT - 1);
      vals[i] = scan_lo + draw(rs, 42, i) % (uint64_t)(scan_hi - scan_lo + 1);
      ops[i] = 3;
    }
  }
  return 0;
}

// Batch sizes floor(2^(lo + (hi - lo) u)) for an uneven stream of batches.
void stm_gen_batch_sizes(size_t count, uint64_t seed, double lo_exp, double hi_exp, uint64_t* out) {
  for (size_t i = 0; i < count; ++i) {
    const double u = unit(draw(seed, 30, i)) - 1e-12;
    out[i] = (uint64_t)std::floor(std::pow(2.0, lo_exp + (hi_exp - lo_exp) * u));
  }
}

// Sorts n keys ascending in place (parallel radix sort); helper for the harnesses.
void stm_sort_u64(uint64_t* keys, size_t n) {
  std::vector<uint64_t> v(keys, keys + n);
  psort_u64(v);
  std::memcpy(keys, v.data(), n * 8);
}

}  // extern "C"
