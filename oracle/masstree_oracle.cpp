// Oracle (test infrastructure only): a host std::map replay of the GPU Masstree
// batch API.  Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark load this library, and only as the checker or the timed CPU
// baseline; the product path never calls it.
//
// Semantics restated from the reference (hyoungjook/IndexinGPU):
//   * key order: lexicographic over u32 slices, a proper prefix first
//     (include/masstree_node_subwarp.hpp:312-314 cmp_key; test/unittest_masstree.cu:153-159
//     std::lexicographical_compare) -> std::vector<uint32_t> operator< is exactly that order.
//   * find: value_lengths[i] = stored length (0 = absent); values get the first
//     min(len, max_value_length) slices; nothing is written for an absent key
//     (include/gpu_masstree.hpp:194-248, include/kernels.hpp:241-256,
//     include/suffix_node_subwarp.hpp:519-544).
//   * insert: insert-if-absent unless update_if_exists (upsert)
//     (include/gpu_masstree.hpp:397-627); a value of length 0 is stored and reads
//     back as length 0, i.e. "absent" to find, but still blocks a later insert
//     (gpu_masstree.hpp:607-617 + :231).
//   * update: overwrite an existing key, false if absent (gpu_masstree.hpp:629-749).
//   * erase: remove, false if absent (gpu_masstree.hpp:751-971).
//   * scan: ascending keys k with lower <= k (inclusive) and, when upper keys are
//     given, k <= upper (inclusive), at most max_count per query; outputs counts,
//     values (min(len,maxV) slices), value lengths, keys and key lengths at strides
//     max_count*maxV / max_count / max_count*maxK (include/kernels.hpp:322-341,
//     include/gpu_masstree.hpp:250-395, masstree_node_subwarp.hpp:315-464).
//   * mixed: request_type insert=0 erase=1 find=2 update=3 (include/macros.hpp:51-58);
//     results[] for writes, values/value_lengths in place for finds
//     (include/kernels.hpp:366-456).  A batch is replayed in index order; the
//     workloads guarantee no key is written twice and no written key is also read
//     in one batch, so any order gives the same outcome.
// Lengths above 65535 slices are rejected by the GPU index (16-bit record
// lengths, include/suffix_node_subwarp.hpp:706); the oracle accepts them.
//
// ORC_PLANT (compile-time, 0 = off) plants one known mistake, so the pinning tests
// can show that they catch it: 1 = find reports the truncated value length,
// 2 = scan reports the truncated key length, 3 = scan reports the truncated value
// length, 4 = find copies one slice fewer than min(len, max_value_length).
// The library the tests and the benchmark load is always built with 0.
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {
using Key = std::vector<uint32_t>;
using Val = std::vector<uint32_t>;
using Map = std::map<Key, Val>;

struct Oracle {
  Map m;
};

inline Key make_key(const uint32_t* keys, uint32_t maxK, const uint32_t* lens, uint64_t i) {
  uint32_t L = lens ? lens[i] : maxK;
  const uint32_t* p = keys + i * (uint64_t)maxK;
  return Key(p, p + L);
}
inline Val make_val(const uint32_t* vals, uint32_t maxV, const uint32_t* lens, uint64_t i) {
  uint32_t L = lens ? lens[i] : maxV;
  const uint32_t* p = vals + i * (uint64_t)maxV;
  return Val(p, p + std::min(L, maxV));
}
inline void write_value(const Val& v, uint32_t* out_vals, uint32_t maxV, uint64_t slot) {
  if (!out_vals) return;
  uint32_t n = std::min<uint32_t>((uint32_t)v.size(), maxV);
  if (ORC_PLANT == 4 && n > 0) n -= 1;
  std::memcpy(out_vals + slot * (uint64_t)maxV, v.data(), n * sizeof(uint32_t));
}
// Lengths are reported in full even when the copy was cut to the output row
// (include/suffix_node_subwarp.hpp:519-544 returns the stored length after copying
// min(len, max) slices; include/gpu_masstree.hpp:227-240 passes it through).
inline uint32_t reported_length(std::size_t stored, uint32_t max_out, bool planted) {
  if (planted) return (uint32_t)std::min<std::size_t>(stored, max_out);
  return (uint32_t)stored;
}
}  // namespace

extern "C" {

void* orc_create() { return new Oracle(); }
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
uint64_t orc_size(void* h) { return static_cast<Oracle*>(h)->m.size(); }
void orc_clear(void* h) { static_cast<Oracle*>(h)->m.clear(); }

// insert (update_if_exists=0: insert-if-absent; 1: upsert). results may be null.
void orc_insert(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* key_lengths,
                const uint32_t* values, uint32_t maxV, const uint32_t* value_lengths,
                uint64_t n, int update_if_exists, uint8_t* results) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t i = 0; i < n; i++) {
    Key k = make_key(keys, maxK, key_lengths, i);
    Val v = make_val(values, maxV, value_lengths, i);
    auto it = m.find(k);
    bool ok;
    if (it == m.end()) {
      m.emplace(std::move(k), std::move(v));
      ok = true;
    } else if (update_if_exists) {
      it->second = std::move(v);
      ok = true;
    } else {
      ok = false;
    }
    if (results) results[i] = ok;
  }
}

void orc_update(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* key_lengths,
                const uint32_t* values, uint32_t maxV, const uint32_t* value_lengths,
                uint64_t n, uint8_t* results) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t i = 0; i < n; i++) {
    auto it = m.find(make_key(keys, maxK, key_lengths, i));
    bool ok = it != m.end();
    if (ok) it->second = make_val(values, maxV, value_lengths, i);
    if (results) results[i] = ok;
  }
}

void orc_erase(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* key_lengths,
               uint64_t n, uint8_t* results) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t i = 0; i < n; i++) {
    bool ok = m.erase(make_key(keys, maxK, key_lengths, i)) > 0;
    if (results) results[i] = ok;
  }
}

void orc_find(void* h, const uint32_t* keys, uint32_t maxK, const uint32_t* key_lengths,
              uint32_t* values, uint32_t maxV, uint32_t* value_lengths, uint64_t n) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t i = 0; i < n; i++) {
    auto it = m.find(make_key(keys, maxK, key_lengths, i));
    if (it == m.end()) {
      value_lengths[i] = 0;
    } else {
      value_lengths[i] = reported_length(it->second.size(), maxV, ORC_PLANT == 1);
      write_value(it->second, values, maxV, i);
    }
  }
}

// scan: per-query limit = limits ? min(limits[q], max_count) : max_count.
void orc_scan(void* h, const uint32_t* lower_keys, const uint32_t* lower_lengths, uint32_t maxK,
              uint32_t max_count, uint64_t nq, const uint32_t* upper_keys,
              const uint32_t* upper_lengths, uint32_t* counts, uint32_t* values, uint32_t maxV,
              uint32_t* value_lengths, uint32_t* out_keys, uint32_t* out_key_lengths,
              const uint32_t* limits) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t q = 0; q < nq; q++) {
    Key lo = make_key(lower_keys, maxK, lower_lengths, q);
    uint32_t limit = max_count;
    if (limits) limit = std::min(limit, limits[q]);
    Key hi;
    if (upper_keys) hi = make_key(upper_keys, maxK, upper_lengths, q);
    uint32_t c = 0;
    for (auto it = m.lower_bound(lo); it != m.end() && c < limit; ++it, ++c) {
      if (upper_keys && hi < it->first) break;
      uint64_t slot = q * (uint64_t)max_count + c;
      if (value_lengths)
        value_lengths[slot] = reported_length(it->second.size(), maxV, ORC_PLANT == 3);
      write_value(it->second, values, maxV, slot);
      if (out_key_lengths)
        out_key_lengths[slot] = reported_length(it->first.size(), maxK, ORC_PLANT == 2);
      if (out_keys) {
        uint32_t nk = std::min<uint32_t>((uint32_t)it->first.size(), maxK);
        std::memcpy(out_keys + slot * (uint64_t)maxK, it->first.data(), nk * sizeof(uint32_t));
      }
    }
    if (counts) counts[q] = c;
  }
}

void orc_mixed(void* h, const uint8_t* types, const uint32_t* keys, uint32_t maxK,
               const uint32_t* key_lengths, uint32_t* values, uint32_t maxV,
               uint32_t* value_lengths, uint8_t* results, uint64_t n) {
  Map& m = static_cast<Oracle*>(h)->m;
  for (uint64_t i = 0; i < n; i++) {
    Key k = make_key(keys, maxK, key_lengths, i);
    switch (types[i]) {
      case 0: {  // insert (insert-if-absent)
        bool ok = m.find(k) == m.end();
        if (ok) m.emplace(std::move(k), make_val(values, maxV, value_lengths, i));
        if (results) results[i] = ok;
        break;
      }
      case 1: {  // erase
        bool ok = m.erase(k) > 0;
        if (results) results[i] = ok;
        break;
      }
      case 2: {  // find
        auto it = m.find(k);
        if (it == m.end()) {
          value_lengths[i] = 0;
        } else {
          value_lengths[i] = reported_length(it->second.size(), maxV, ORC_PLANT == 1);
          write_value(it->second, values, maxV, i);
        }
        break;
      }
      case 3: {  // update
        auto it = m.find(k);
        bool ok = it != m.end();
        if (ok) it->second = make_val(values, maxV, value_lengths, i);
        if (results) results[i] = ok;
        break;
      }
      default:
        break;
    }
  }
}

// Export the whole map in key order.  Call with null buffers first to get the
// totals (returns entry count; *key_slices / *value_slices get slice totals).
uint64_t orc_dump(void* h, uint32_t* key_offsets, uint32_t* key_data, uint32_t* value_offsets,
                  uint32_t* value_data, uint64_t* key_slices, uint64_t* value_slices) {
  Map& m = static_cast<Oracle*>(h)->m;
  uint64_t ks = 0, vs = 0, i = 0;
  for (auto& kv : m) {
    if (key_offsets) key_offsets[i] = (uint32_t)ks;
    if (value_offsets) value_offsets[i] = (uint32_t)vs;
    if (key_data) std::memcpy(key_data + ks, kv.first.data(), kv.first.size() * 4);
    if (value_data) std::memcpy(value_data + vs, kv.second.data(), kv.second.size() * 4);
    ks += kv.first.size();
    vs += kv.second.size();
    i++;
  }
  if (key_offsets) key_offsets[i] = (uint32_t)ks;
  if (value_offsets) value_offsets[i] = (uint32_t)vs;
  if (key_slices) *key_slices = ks;
  if (value_slices) *value_slices = vs;
  return i;
}

}  // extern "C"
