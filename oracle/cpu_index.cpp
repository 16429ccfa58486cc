// Multi-threaded CPU ordered index: the timed CPU baseline beside the GPU index, and a
// second checker.  Test and benchmark infrastructure only (the product never calls it).
//
// Shape: the reference's CPU adapter (evaluation/cpu_masstree_adapter.hpp:36-131 in
// hyoungjook/IndexinGPU) -- one index shared by worker threads, each calling
// insert / update / erase / find / scan for one key at a time -- and its driver's
// round-robin striding of a batch over std::thread::hardware_concurrency() workers
// (evaluation/generate_workload.hpp:38-52).  The reference's adapter wraps
// masstree-beta, which is not in this repository; this index is NOT masstree-beta: it
// is std::map shards, split by the top bits of the first key slice (range partitioning
// keeps the global key order, so a scan walks the shards in order), each behind a
// reader/writer lock.  Key order and value semantics are the oracle's (lexicographic
// u32 slices, proper prefix first; value lengths with 0 = absent on find).
//
// Conventions of the reference's CPU adapter kept here:
//   * keys are stored as byte strings, each u32 slice byte-swapped to big-endian, so that
//     byte order equals slice order (include/generate_workload.hpp:291-297 and
//     evaluation/universal_bench.cu:136-140 swap the slices before handing them to
//     masstree-beta); a proper prefix sorts first, as in the GPU index;
//   * the adapter-shaped entry points (cpx_put1 / cpx_get1 / cpx_scan) take one-slice
//     values, report an absent key as 0xFFFFFFFF (cpu_masstree_adapter.hpp:42,120-125)
//     and insert by overwriting an existing value (cpu_masstree_adapter.hpp:96-104).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Key = std::string;  // big-endian bytes of the u32 slices
using Val = std::vector<uint32_t>;
constexpr uint32_t kAbsent = 0xFFFFFFFFu;

struct shard {
  std::shared_mutex mu;
  std::map<Key, Val> m;
};

struct cpu_index {
  uint32_t bits;
  std::vector<std::unique_ptr<shard>> shards;
  explicit cpu_index(uint32_t b) : bits(b) {
    for (uint32_t i = 0; i < (1u << b); i++) shards.emplace_back(new shard());
  }
  // shard of a key: the top `bits` bits of its first slice (its first bytes)
  uint32_t shard_of(const Key& k) const {
    if (!bits || k.empty()) return 0;
    uint32_t first = 0;
    for (size_t b = 0; b < 4; b++)
      first = (first << 8) | (b < k.size() ? (uint8_t)k[b] : 0u);
    return first >> (32 - bits);
  }
  shard& of(const Key& k) { return *shards[shard_of(k)]; }
};

Key make_key(const uint32_t* keys, uint32_t maxK, const uint32_t* lens, uint64_t i) {
  const uint32_t L = lens ? lens[i] : maxK;
  const uint32_t* p = keys + i * (uint64_t)maxK;
  Key k(4ull * L, '\0');
  for (uint32_t j = 0; j < L; j++) {
    const uint32_t v = __builtin_bswap32(p[j]);
    std::memcpy(&k[4ull * j], &v, 4);
  }
  return k;
}

// runs f(i) for i in [0, n) on `threads` workers, worker t taking i = t, t + threads, ...
template <class F>
void striped(uint64_t n, uint32_t threads, F&& f) {
  threads = std::max<uint32_t>(1, threads);
  if (threads == 1) {
    for (uint64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      for (uint64_t i = t; i < n; i += threads) f(i);
    });
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

void* cpx_create(uint32_t shard_bits) { return new cpu_index(std::min<uint32_t>(shard_bits, 16)); }
void cpx_destroy(void* h) { delete static_cast<cpu_index*>(h); }

uint64_t cpx_size(void* h) {
  auto* x = static_cast<cpu_index*>(h);
  uint64_t n = 0;
  for (auto& s : x->shards) n += s->m.size();
  return n;
}

// insert (insert-if-absent, or upsert when update_if_exists); results nullable
void cpx_insert(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* lens,
                const uint32_t* values, uint32_t maxV, const uint32_t* vlens, uint64_t n,
                int update_if_exists, uint8_t* results, uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(n, threads, [&](uint64_t i) {
    Key k = make_key(keys, maxK, lens, i);
    const uint32_t vl = vlens ? vlens[i] : maxV;
    const uint32_t* vp = values + i * (uint64_t)maxV;
    shard& s = x->of(k);
    std::unique_lock<std::shared_mutex> g(s.mu);
    auto it = s.m.find(k);
    bool ok = it == s.m.end();
    if (ok)
      s.m.emplace(std::move(k), Val(vp, vp + vl));
    else if (update_if_exists)
      it->second.assign(vp, vp + vl);
    if (results) results[i] = ok || update_if_exists;
  });
}

void cpx_erase(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* lens, uint64_t n,
               uint8_t* results, uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(n, threads, [&](uint64_t i) {
    Key k = make_key(keys, maxK, lens, i);
    shard& s = x->of(k);
    std::unique_lock<std::shared_mutex> g(s.mu);
    const bool ok = s.m.erase(k) > 0;
    if (results) results[i] = ok;
  });
}

// find: out_lens[i] = stored length (0 = absent); out_values get min(len, maxV) slices
void cpx_find(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* lens,
              uint32_t* out_values, uint32_t maxV, uint32_t* out_lens, uint64_t n,
              uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(n, threads, [&](uint64_t i) {
    Key k = make_key(keys, maxK, lens, i);
    shard& s = x->of(k);
    std::shared_lock<std::shared_mutex> g(s.mu);
    auto it = s.m.find(k);
    if (it == s.m.end()) {
      out_lens[i] = 0;
      return;
    }
    const Val& v = it->second;
    out_lens[i] = (uint32_t)v.size();
    const uint32_t c = std::min<uint32_t>((uint32_t)v.size(), maxV);
    if (c) std::memcpy(out_values + i * (uint64_t)maxV, v.data(), 4ull * c);
  });
}

// scan: the adapter's shape -- up to `count` values (first slice of each) of keys >= the
// lower key in order, the rest of the row filled with 0xFFFFFFFF; counts[q] = results
void cpx_scan(void* h, const uint32_t* lower, uint32_t maxK, const uint32_t* lens, uint32_t count,
              uint64_t nq, uint32_t* out_values, uint32_t* counts, uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(nq, threads, [&](uint64_t q) {
    Key k = make_key(lower, maxK, lens, q);
    uint32_t* row = out_values + q * (uint64_t)count;
    uint32_t c = 0;
    const uint32_t first = x->shard_of(k);
    for (uint32_t sh = first; sh < x->shards.size() && c < count; sh++) {
      shard& s = *x->shards[sh];
      std::shared_lock<std::shared_mutex> g(s.mu);
      for (auto it = s.m.lower_bound(k); it != s.m.end() && c < count; ++it)
        row[c++] = it->second.empty() ? kAbsent : it->second[0];
    }
    if (counts) counts[q] = c;
    for (uint32_t j = c; j < count; j++) row[j] = kAbsent;
  });
}

// The adapter's insert (cpu_masstree_adapter.hpp:96-104): one-slice value, an existing
// key's value is overwritten.
void cpx_put1(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* lens,
              const uint32_t* values, uint64_t n, uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(n, threads, [&](uint64_t i) {
    Key k = make_key(keys, maxK, lens, i);
    shard& s = x->of(k);
    std::unique_lock<std::shared_mutex> g(s.mu);
    s.m[std::move(k)].assign(1, values[i]);
  });
}

// The adapter's find (cpu_masstree_adapter.hpp:120-125): the first value slice, or
// 0xFFFFFFFF for an absent key.
void cpx_get1(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* lens,
              uint32_t* out, uint64_t n, uint32_t threads) {
  auto* x = static_cast<cpu_index*>(h);
  striped(n, threads, [&](uint64_t i) {
    Key k = make_key(keys, maxK, lens, i);
    shard& s = x->of(k);
    std::shared_lock<std::shared_mutex> g(s.mu);
    auto it = s.m.find(k);
    out[i] = (it == s.m.end() || it->second.empty()) ? kAbsent : it->second[0];
  });
}

}  // extern "C"
