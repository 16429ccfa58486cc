// C ABI over GpuMasstree::gpu_masstree (declared in include/gpu_masstree_c.h).
// Instantiates the index for tile widths 16, 32 and 64 and dispatches at run time.
#include <cstring>
#include <memory>
#include <string>

#include "gpu_masstree_c.h"
#include "indexingpu_amd/gpu_masstree.hpp"

namespace {

thread_local std::string g_last_error;

using alloc_t = indexingpu_amd::simple_slab_allocator<128>;
using reclaim_t = indexingpu_amd::simple_debra_reclaimer<>;

struct index_base {
  virtual ~index_base() = default;
  virtual uint32_t tile() const = 0;
  virtual void find(const uint32_t*, uint32_t, const uint32_t*, uint32_t*, uint32_t, uint32_t*,
                    uint32_t, bool, hipStream_t) = 0;
  virtual void insert(const uint32_t*, uint32_t, const uint32_t*, const uint32_t*, uint32_t,
                      const uint32_t*, uint32_t, bool, bool*, hipStream_t) = 0;
  virtual void update(const uint32_t*, uint32_t, const uint32_t*, const uint32_t*, uint32_t,
                      const uint32_t*, uint32_t, bool*, hipStream_t) = 0;
  virtual void erase(const uint32_t*, uint32_t, const uint32_t*, uint32_t, bool*, hipStream_t) = 0;
  virtual void scan(const uint32_t*, const uint32_t*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                    const uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t*, uint32_t*,
                    uint32_t*, const uint32_t*, hipStream_t) = 0;
  virtual void mixed(const uint8_t*, const uint32_t*, uint32_t, const uint32_t*, uint32_t*,
                     uint32_t, uint32_t*, bool*, uint32_t, hipStream_t) = 0;
  virtual uint32_t device_error(bool clear) = 0;
  virtual indexingpu_amd::walk_stats walk(const indexingpu_amd::entry_emit_fn&) = 0;
  virtual uint64_t slabs_allocated() = 0;
  virtual uint64_t pool_slabs() = 0;
  virtual void set_find_depth(uint32_t k) = 0;
};

template <uint32_t T>
struct index_impl final : index_base {
  std::unique_ptr<alloc_t> alloc;
  std::unique_ptr<reclaim_t> reclaim;
  std::unique_ptr<GpuMasstree::gpu_masstree<alloc_t, reclaim_t, T>> tree;

  index_impl(uint64_t pool_bytes, float pool_ratio, uint32_t retire_capacity) {
    if (pool_bytes)
      alloc.reset(new alloc_t(indexingpu_amd::pool_bytes{pool_bytes}));
    else
      alloc.reset(new alloc_t(pool_ratio > 0.f ? pool_ratio : 0.9f));
    reclaim.reset(retire_capacity ? new reclaim_t(retire_capacity) : new reclaim_t());
    tree.reset(new GpuMasstree::gpu_masstree<alloc_t, reclaim_t, T>(*alloc, *reclaim));
  }
  ~index_impl() override {
    (void)hipDeviceSynchronize();
    tree.reset();
    reclaim.reset();
    alloc.reset();
  }
  uint32_t tile() const override { return T; }
  void find(const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t* v, uint32_t mv,
            uint32_t* vl, uint32_t n, bool conc, hipStream_t s) override {
    if (conc)
      tree->template find<true>(k, mk, kl, v, mv, vl, n, s);
    else
      tree->template find<false>(k, mk, kl, v, mv, vl, n, s);
  }
  void insert(const uint32_t* k, uint32_t mk, const uint32_t* kl, const uint32_t* v, uint32_t mv,
              const uint32_t* vl, uint32_t n, bool upsert, bool* r, hipStream_t s) override {
    if (upsert)
      tree->template insert<true>(k, mk, kl, v, mv, vl, n, s, r);
    else
      tree->template insert<false>(k, mk, kl, v, mv, vl, n, s, r);
  }
  void update(const uint32_t* k, uint32_t mk, const uint32_t* kl, const uint32_t* v, uint32_t mv,
              const uint32_t* vl, uint32_t n, bool* r, hipStream_t s) override {
    tree->update(k, mk, kl, v, mv, vl, n, s, r);
  }
  void erase(const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t n, bool* r,
             hipStream_t s) override {
    tree->erase(k, mk, kl, n, s, r);
  }
  void scan(const uint32_t* lk, const uint32_t* ll, uint32_t mk, uint32_t mc, uint32_t nq,
            const uint32_t* uk, const uint32_t* ul, uint32_t* counts, uint32_t* vals, uint32_t mv,
            uint32_t* vlens, uint32_t* ok, uint32_t* okl, const uint32_t* lim,
            hipStream_t s) override {
    if (uk)
      tree->template scan<true>(lk, ll, mk, mc, nq, uk, ul, counts, vals, mv, vlens, ok, okl, s, lim);
    else
      tree->template scan<false>(lk, ll, mk, mc, nq, nullptr, nullptr, counts, vals, mv, vlens, ok,
                                 okl, s, lim);
  }
  void mixed(const uint8_t* t, const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t* v,
             uint32_t mv, uint32_t* vl, bool* r, uint32_t n, hipStream_t s) override {
    tree->mixed_batch(reinterpret_cast<const kernels::request_type*>(t), k, mk, kl, v, mv, vl, r,
                      n, s);
  }
  uint32_t device_error(bool clear) override {
    IGPU_HIP_TRY(hipDeviceSynchronize());
    uint32_t h = 0;
    IGPU_HIP_TRY(hipMemcpy(&h, tree->err_, 4, hipMemcpyDeviceToHost));
    if (clear) IGPU_HIP_TRY(hipMemset(tree->err_, 0, 4));
    return h;
  }
  indexingpu_amd::walk_stats walk(const indexingpu_amd::entry_emit_fn& f) override {
    return tree->walk(f);
  }
  uint64_t slabs_allocated() override { return alloc->num_allocated_slabs(); }
  uint64_t pool_slabs() override { return alloc->total_slabs(); }
  void set_find_depth(uint32_t k) override { tree->set_find_depth(k); }
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      g_last_error = std::string("HIP error: ") + hipGetErrorString(e);
      return GMT_EHIP;
    }
    return GMT_OK;
  } catch (const std::exception& ex) {
    g_last_error = ex.what();
    return GMT_EHIP;
  }
}

inline index_base* B(gmt_index* h) { return reinterpret_cast<index_base*>(h); }
inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int bad(const char* msg) {
  g_last_error = msg;
  return GMT_EINVAL;
}

}  // namespace

extern "C" {

int gmt_create(uint32_t tile_size, uint64_t pool_bytes, float pool_ratio,
               uint32_t retire_capacity, gmt_index** out) {
  if (!out) return bad("gmt_create: null output handle");
  *out = nullptr;
  index_base* p = nullptr;
  int rc = guarded([&] {
    switch (tile_size) {
      case 16: p = new index_impl<16>(pool_bytes, pool_ratio, retire_capacity); break;
      case 32: p = new index_impl<32>(pool_bytes, pool_ratio, retire_capacity); break;
      case 64: p = new index_impl<64>(pool_bytes, pool_ratio, retire_capacity); break;
      default: break;
    }
  });
  if (rc != GMT_OK) return rc;
  if (!p) return bad("gmt_create: tile_size must be 16, 32 or 64");
  *out = reinterpret_cast<gmt_index*>(p);
  return GMT_OK;
}

int gmt_destroy(gmt_index* index) {
  if (!index) return bad("gmt_destroy: null handle");
  return guarded([&] { delete B(index); });
}

int gmt_find(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
             const uint32_t* key_lengths, uint32_t* values, uint32_t max_value_length,
             uint32_t* value_lengths, uint32_t num_keys, int concurrent, void* stream) {
  if (!index) return bad("gmt_find: null handle");
  if (num_keys && (!keys || !value_lengths)) return bad("gmt_find: null keys or value_lengths");
  if (num_keys && max_value_length && !values) return bad("gmt_find: null values");
  return guarded([&] {
    B(index)->find(keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
                   num_keys, concurrent != 0, S(stream));
  });
}

int gmt_insert(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, int update_if_exists,
               uint8_t* results, void* stream) {
  if (!index) return bad("gmt_insert: null handle");
  if (num_keys && (!keys || (!values && max_value_length)))
    return bad("gmt_insert: null keys or values");
  return guarded([&] {
    B(index)->insert(keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
                     num_keys, update_if_exists != 0, reinterpret_cast<bool*>(results), S(stream));
  });
}

int gmt_update(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, uint8_t* results, void* stream) {
  if (!index) return bad("gmt_update: null handle");
  if (num_keys && (!keys || (!values && max_value_length)))
    return bad("gmt_update: null keys or values");
  return guarded([&] {
    B(index)->update(keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
                     num_keys, reinterpret_cast<bool*>(results), S(stream));
  });
}

int gmt_erase(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
              const uint32_t* key_lengths, uint32_t num_keys, uint8_t* results, void* stream) {
  if (!index) return bad("gmt_erase: null handle");
  if (num_keys && !keys) return bad("gmt_erase: null keys");
  return guarded([&] {
    B(index)->erase(keys, max_key_length, key_lengths, num_keys, reinterpret_cast<bool*>(results),
                    S(stream));
  });
}

int gmt_scan(gmt_index* index, const uint32_t* lower_keys, const uint32_t* lower_key_lengths,
             uint32_t max_key_length, uint32_t max_count_per_query, uint32_t num_queries,
             const uint32_t* upper_keys, const uint32_t* upper_key_lengths, uint32_t* counts,
             uint32_t* values, uint32_t max_value_length, uint32_t* value_lengths,
             uint32_t* out_keys, uint32_t* out_key_lengths, const uint32_t* per_query_limits,
             void* stream) {
  if (!index) return bad("gmt_scan: null handle");
  if (num_queries && !lower_keys) return bad("gmt_scan: null lower_keys");
  return guarded([&] {
    B(index)->scan(lower_keys, lower_key_lengths, max_key_length, max_count_per_query, num_queries,
                   upper_keys, upper_key_lengths, counts, values, max_value_length, value_lengths,
                   out_keys, out_key_lengths, per_query_limits, S(stream));
  });
}

int gmt_mixed_batch(gmt_index* index, const uint8_t* types, const uint32_t* keys,
                    uint32_t max_key_length, const uint32_t* key_lengths, uint32_t* values,
                    uint32_t max_value_length, uint32_t* value_lengths, uint8_t* results,
                    uint32_t num_requests, void* stream) {
  if (!index) return bad("gmt_mixed_batch: null handle");
  if (num_requests && (!types || !keys || !value_lengths))
    return bad("gmt_mixed_batch: null types, keys or value_lengths");
  return guarded([&] {
    B(index)->mixed(types, keys, max_key_length, key_lengths, values, max_value_length,
                    value_lengths, reinterpret_cast<bool*>(results), num_requests, S(stream));
  });
}

uint32_t gmt_device_error(gmt_index* index, int clear) {
  if (!index) return 0xFFFFFFFFu;
  uint32_t e = 0xFFFFFFFFu;
  if (guarded([&] { e = B(index)->device_error(clear != 0); }) != GMT_OK) return 0xFFFFFFFFu;
  return e;
}

int gmt_validate(gmt_index* index, uint64_t* entries, uint64_t* nodes, uint64_t* record_slabs,
                 char* msg, uint32_t msg_len) {
  if (!index) return bad("gmt_validate: null handle");
  indexingpu_amd::walk_stats s;
  int rc = guarded([&] { s = B(index)->walk(nullptr); });
  if (rc != GMT_OK) return -rc;
  if (entries) *entries = s.entries;
  if (nodes) *nodes = s.nodes;
  if (record_slabs) *record_slabs = s.record_slabs;
  if (msg && msg_len) {
    std::strncpy(msg, s.error.c_str(), msg_len - 1);
    msg[msg_len - 1] = 0;
  }
  return s.error.empty() ? 1 : 0;
}

int64_t gmt_dump(gmt_index* index, uint32_t* key_offsets, uint32_t* key_data,
                 uint32_t* value_offsets, uint32_t* value_data, uint64_t* key_slices,
                 uint64_t* value_slices) {
  if (!index) return -1;
  uint64_t ks = 0, vs = 0, i = 0;
  indexingpu_amd::walk_stats s;
  int rc = guarded([&] {
    s = B(index)->walk([&](const std::vector<uint32_t>& k, const std::vector<uint32_t>& v) {
      if (key_offsets) key_offsets[i] = (uint32_t)ks;
      if (value_offsets) value_offsets[i] = (uint32_t)vs;
      if (key_data) std::memcpy(key_data + ks, k.data(), k.size() * 4);
      if (value_data) std::memcpy(value_data + vs, v.data(), v.size() * 4);
      ks += k.size();
      vs += v.size();
      i++;
    });
  });
  if (rc != GMT_OK) return -1;
  if (!s.error.empty()) {
    g_last_error = "gmt_dump: " + s.error;
    return -1;
  }
  if (key_offsets) key_offsets[i] = (uint32_t)ks;
  if (value_offsets) value_offsets[i] = (uint32_t)vs;
  if (key_slices) *key_slices = ks;
  if (value_slices) *value_slices = vs;
  return (int64_t)i;
}

uint64_t gmt_slabs_allocated(gmt_index* index) {
  if (!index) return 0;
  uint64_t v = 0;
  guarded([&] { v = B(index)->slabs_allocated(); });
  return v;
}

uint64_t gmt_pool_slabs(gmt_index* index) { return index ? B(index)->pool_slabs() : 0; }

int gmt_set_find_depth(gmt_index* index, uint32_t depth) {
  if (!index) return bad("gmt_set_find_depth: null handle");
  B(index)->set_find_depth(depth);
  return GMT_OK;
}

uint32_t gmt_tile_size(gmt_index* index) { return index ? B(index)->tile() : 0; }

const char* gmt_last_error_message(void) { return g_last_error.c_str(); }

}  // extern "C"
