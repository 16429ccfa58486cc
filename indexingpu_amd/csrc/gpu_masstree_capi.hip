// C ABI over GpuMasstree::gpu_masstree (declared in include/gpu_masstree_c.h).
// One index holds the tree for tile widths 16, 32 and 64 at once (the 16-lane instance
// owns it, the others attach to the same state: the node layout does not depend on the
// tile width), so the active width is a run-time choice.
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "gpu_masstree_c.h"
#include "indexingpu_amd/gpu_masstree.hpp"
#include "indexingpu_amd/server.hpp"

namespace {

thread_local std::string g_last_error;

using alloc_t = indexingpu_amd::simple_slab_allocator<128>;
using reclaim_t = indexingpu_amd::simple_debra_reclaimer<>;
template <uint32_t T>
using tree_t = GpuMasstree::gpu_masstree<alloc_t, reclaim_t, T>;

struct gmt_impl {
  std::unique_ptr<alloc_t> alloc;
  std::unique_ptr<reclaim_t> reclaim;
  std::unique_ptr<tree_t<16>> t16;
  std::unique_ptr<tree_t<32>> t32;
  std::unique_ptr<tree_t<64>> t64;
  uint32_t tile = 16;

  gmt_impl(uint32_t tile_size, uint64_t pool_bytes, float pool_ratio, uint32_t retire_capacity)
      : tile(tile_size) {
    if (pool_bytes)
      alloc.reset(new alloc_t(indexingpu_amd::pool_bytes{pool_bytes}));
    else
      alloc.reset(new alloc_t(pool_ratio > 0.f ? pool_ratio : 0.9f));
    reclaim.reset(retire_capacity ? new reclaim_t(retire_capacity) : new reclaim_t());
    t16.reset(new tree_t<16>(*alloc, *reclaim));
    t32.reset(new tree_t<32>(*t16));
    t64.reset(new tree_t<64>(*t16));
  }
  ~gmt_impl() {
    (void)hipDeviceSynchronize();
    t64.reset();
    t32.reset();
    t16.reset();
    reclaim.reset();
    alloc.reset();
  }
  // runs f on the instance of the active tile width
  template <class F>
  void with_tree(F&& f) {
    switch (tile) {
      case 32: f(*t32); break;
      case 64: f(*t64); break;
      default: f(*t16); break;
    }
  }
  // settings that live per instance are applied to all three
  template <class F>
  void on_all(F&& f) {
    f(*t16);
    f(*t32);
    f(*t64);
  }
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

inline gmt_impl* B(gmt_index* h) { return reinterpret_cast<gmt_impl*>(h); }

// Scratch of gmt_find_noalloc: the two-pass find's per-key word for up to
// GMT_NOALLOC_MAX_KEYS keys, a fixed device array (each GPU the library runs on holds its
// own copy until the process ends).  Host code never names the array itself in a
// kernel argument or copy: its device address, one per device, comes from
// hipGetSymbolAddress at set-up (gmt_create on that device, or the first
// gmt_find_noalloc there).
static __device__ __align__(256) uint32_t scratch[GMT_NOALLOC_MAX_KEYS];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (set up once per device); null on
// failure, with the error message set.
uint32_t* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    g_last_error = "gmt_find_noalloc: no usable current device";
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_last_error = std::string("gmt_find_noalloc: hipGetSymbolAddress failed: ") +
                     hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  return g_scratch_ptr[dev];
}
inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int bad(const char* msg) {
  g_last_error = msg;
  return GMT_EINVAL;
}

bool valid_tile(uint32_t t) { return t == 16 || t == 32 || t == 64; }

}  // namespace

extern "C" {

int gmt_create(uint32_t tile_size, uint64_t pool_bytes, float pool_ratio,
               uint32_t retire_capacity, gmt_index** out) {
  if (!out) return bad("gmt_create: null output handle");
  *out = nullptr;
  if (!valid_tile(tile_size)) return bad("gmt_create: tile_size must be 16, 32 or 64");
  gmt_impl* p = nullptr;
  int rc = guarded([&] {
    p = new gmt_impl(tile_size, pool_bytes, pool_ratio, retire_capacity);
    // set-up of gmt_find_noalloc on this device: its scratch address and the launch
    // parameters of the find kernels, so that the entry point itself allocates nothing
    if (!noalloc_scratch()) throw std::runtime_error(g_last_error);
    p->on_all([](auto& t) { t.warm_find_launches(); });
  });
  if (rc != GMT_OK) {
    delete p;
    return rc;
  }
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
    B(index)->with_tree([&](auto& t) {
      if (concurrent)
        t.template find<true>(keys, max_key_length, key_lengths, values, max_value_length,
                              value_lengths, num_keys, S(stream));
      else
        t.template find<false>(keys, max_key_length, key_lengths, values, max_value_length,
                               value_lengths, num_keys, S(stream));
    });
  });
}

int gmt_find_noalloc(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
                     const uint32_t* key_lengths, uint32_t* values, uint32_t max_value_length,
                     uint32_t* value_lengths, uint32_t num_keys, void* stream) {
  if (!index) return bad("gmt_find_noalloc: null handle");
  if (num_keys && (!keys || !value_lengths)) return bad("gmt_find_noalloc: null keys or value_lengths");
  if (num_keys && max_value_length && !values) return bad("gmt_find_noalloc: null values");
  if (num_keys > GMT_NOALLOC_MAX_KEYS)
    return bad("gmt_find_noalloc: num_keys above GMT_NOALLOC_MAX_KEYS");
  if (max_key_length > 65536) return bad("gmt_find_noalloc: max_key_length above 65536");
  uint32_t* sp = noalloc_scratch();
  if (!sp) return GMT_EHIP;
  return guarded([&] {
    B(index)->with_tree([&](auto& t) {
      t.find_with_scratch(keys, max_key_length, key_lengths, values, max_value_length,
                          value_lengths, num_keys, sp, GMT_NOALLOC_MAX_KEYS, S(stream));
    });
  });
}

int gmt_insert(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, int update_if_exists,
               uint8_t* results, void* stream) {
  if (!index) return bad("gmt_insert: null handle");
  if (num_keys && (!keys || (!values && max_value_length)))
    return bad("gmt_insert: null keys or values");
  bool* r = reinterpret_cast<bool*>(results);
  return guarded([&] {
    B(index)->with_tree([&](auto& t) {
      if (update_if_exists)
        t.template insert<true>(keys, max_key_length, key_lengths, values, max_value_length,
                                value_lengths, num_keys, S(stream), r);
      else
        t.template insert<false>(keys, max_key_length, key_lengths, values, max_value_length,
                                 value_lengths, num_keys, S(stream), r);
    });
  });
}

int gmt_update(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, uint8_t* results, void* stream) {
  if (!index) return bad("gmt_update: null handle");
  if (num_keys && (!keys || (!values && max_value_length)))
    return bad("gmt_update: null keys or values");
  bool* r = reinterpret_cast<bool*>(results);
  return guarded([&] {
    B(index)->with_tree([&](auto& t) {
      t.update(keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
               num_keys, S(stream), r);
    });
  });
}

int gmt_erase(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
              const uint32_t* key_lengths, uint32_t num_keys, uint8_t* results, void* stream) {
  if (!index) return bad("gmt_erase: null handle");
  if (num_keys && !keys) return bad("gmt_erase: null keys");
  bool* r = reinterpret_cast<bool*>(results);
  return guarded([&] {
    B(index)->with_tree(
        [&](auto& t) { t.erase(keys, max_key_length, key_lengths, num_keys, S(stream), r); });
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
    B(index)->with_tree([&](auto& t) {
      if (upper_keys)
        t.template scan<true>(lower_keys, lower_key_lengths, max_key_length, max_count_per_query,
                              num_queries, upper_keys, upper_key_lengths, counts, values,
                              max_value_length, value_lengths, out_keys, out_key_lengths,
                              S(stream), per_query_limits);
      else
        t.template scan<false>(lower_keys, lower_key_lengths, max_key_length,
                               max_count_per_query, num_queries, nullptr, nullptr, counts, values,
                               max_value_length, value_lengths, out_keys, out_key_lengths,
                               S(stream), per_query_limits);
    });
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
    B(index)->with_tree([&](auto& t) {
      t.mixed_batch(reinterpret_cast<const kernels::request_type*>(types), keys, max_key_length,
                    key_lengths, values, max_value_length, value_lengths,
                    reinterpret_cast<bool*>(results), num_requests, S(stream));
    });
  });
}

int gmt_multi_batch(gmt_index* index, const uint8_t* types, const uint32_t* keys,
                    uint32_t max_key_length, const uint32_t* key_lengths, uint32_t* values,
                    uint32_t max_value_length, uint32_t* value_lengths, uint8_t* results,
                    const uint64_t* offsets, uint32_t num_batches, uint64_t total_requests,
                    void* stream) {
  if (!index) return bad("gmt_multi_batch: null handle");
  if (num_batches && (!types || !keys || !value_lengths || !offsets))
    return bad("gmt_multi_batch: null types, keys, value_lengths or offsets");
  return guarded([&] {
    B(index)->with_tree([&](auto& t) {
      t.multi_batch(reinterpret_cast<const kernels::request_type*>(types), keys, max_key_length,
                    key_lengths, values, max_value_length, value_lengths,
                    reinterpret_cast<bool*>(results), offsets, num_batches, total_requests,
                    S(stream));
    });
  });
}

uint32_t gmt_device_error(gmt_index* index, int clear) {
  if (!index) return 0xFFFFFFFFu;
  uint32_t e = 0xFFFFFFFFu;
  int rc = guarded([&] {
    uint32_t* d = B(index)->t16->err_;
    IGPU_HIP_TRY(hipDeviceSynchronize());
    IGPU_HIP_TRY(hipMemcpy(&e, d, 4, hipMemcpyDeviceToHost));
    if (clear) B(index)->t16->clear_error();
  });
  return rc == GMT_OK ? e : 0xFFFFFFFFu;
}

int gmt_guard_trip(gmt_index* index, uint32_t* out) {
  if (!index || !out) return -bad("gmt_guard_trip: null argument");
  bool present = false;
  int rc = guarded([&] { present = B(index)->t16->last_guard_trip(out); });
  if (rc != GMT_OK) return -rc;
  return present ? 1 : 0;
}

int gmt_validate(gmt_index* index, uint64_t* entries, uint64_t* nodes, uint64_t* record_slabs,
                 char* msg, uint32_t msg_len) {
  if (!index) return bad("gmt_validate: null handle");
  indexingpu_amd::walk_stats s;
  int rc = guarded([&] { s = B(index)->t16->walk(nullptr); });
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
    s = B(index)->t16->walk([&](const std::vector<uint32_t>& k, const std::vector<uint32_t>& v) {
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
  guarded([&] { v = B(index)->alloc->num_allocated_slabs(); });
  return v;
}

uint64_t gmt_pool_slabs(gmt_index* index) { return index ? B(index)->alloc->total_slabs() : 0; }

int gmt_set_multi_grid_percent(gmt_index* index, uint32_t percent) {
  if (!index) return bad("gmt_set_multi_grid_percent: null handle");
  B(index)->on_all([&](auto& t) { t.set_multi_grid_percent(percent); });
  return GMT_OK;
}

int gmt_set_find_depth(gmt_index* index, uint32_t depth) {
  if (!index) return bad("gmt_set_find_depth: null handle");
  B(index)->on_all([&](auto& t) { t.set_find_depth(depth); });
  return GMT_OK;
}

int gmt_set_find_passes(gmt_index* index, uint32_t passes) {
  if (!index) return bad("gmt_set_find_passes: null handle");
  if (passes != 1 && passes != 2) return bad("gmt_set_find_passes: passes must be 1 or 2");
  B(index)->on_all([&](auto& t) { t.set_find_passes(passes); });
  return GMT_OK;
}

int gmt_set_record_tile(gmt_index* index, uint32_t record_tile) {
  if (!index) return bad("gmt_set_record_tile: null handle");
  if (record_tile != 0 && record_tile != 1 && record_tile != 2 && record_tile != 4 &&
      record_tile != 8 && record_tile != 16)
    return bad("gmt_set_record_tile: record tile must be 0, 1, 2, 4, 8 or 16");
  B(index)->on_all([&](auto& t) { t.set_record_tile(record_tile); });
  return GMT_OK;
}

int gmt_reserve_find_scratch(gmt_index* index, uint64_t max_requests) {
  if (!index) return bad("gmt_reserve_find_scratch: null handle");
  // every tile width's instance keeps its own scratch
  return guarded([&] { B(index)->on_all([&](auto& t) { t.reserve_find_scratch(max_requests); }); });
}

int gmt_set_write_hints(gmt_index* index, uint32_t min_requests) {
  if (!index) return bad("gmt_set_write_hints: null handle");
  B(index)->on_all([&](auto& t) { t.set_write_hints(min_requests); });
  return GMT_OK;
}

int gmt_set_mixed_split(gmt_index* index, uint32_t min_requests) {
  if (!index) return bad("gmt_set_mixed_split: null handle");
  B(index)->on_all([&](auto& t) { t.set_mixed_split(min_requests); });
  return GMT_OK;
}

int gmt_set_multi_mode(gmt_index* index, uint32_t mode) {
  if (!index) return bad("gmt_set_multi_mode: null handle");
  if (mode > 1) return bad("gmt_set_multi_mode: mode must be 0 or 1");
  B(index)->on_all([&](auto& t) { t.set_multi_mode(mode); });
  return GMT_OK;
}

int gmt_set_find_tile(gmt_index* index, uint32_t find_tile) {
  if (!index) return bad("gmt_set_find_tile: null handle");
  if (find_tile > 16 || (find_tile & (find_tile - 1)) != 0)
    return bad("gmt_set_find_tile: find_tile must be 0, 1, 2, 4, 8 or 16");
  B(index)->on_all([&](auto& t) { t.set_find_tile(find_tile); });
  return GMT_OK;
}

int gmt_set_find_shortcut(gmt_index* index, int on) {
  if (!index) return bad("gmt_set_find_shortcut: null handle");
  return guarded([&] { B(index)->on_all([&](auto& t) { t.set_find_shortcut(on != 0); }); });
}

uint32_t gmt_find_shortcut_bits(gmt_index* index) {
  if (!index) return 0;
  uint32_t b = 0;
  guarded([&] { B(index)->on_all([&](auto& t) { b = std::max(b, t.find_shortcut_bits()); }); });
  return b;
}

int gmt_set_fanout(gmt_index* index, uint32_t border_cap, uint32_t interior_cap) {
  if (!index) return bad("gmt_set_fanout: null handle");
  if ((border_cap != 0 && (border_cap < 2 || border_cap > indexingpu_amd::kBorderCap)) ||
      (interior_cap != 0 && (interior_cap < 3 || interior_cap > indexingpu_amd::kInteriorCap)))
    return bad("gmt_set_fanout: border_cap must be 0 or 2..14, interior_cap 0 or 3..15");
  B(index)->on_all([&](auto& t) { t.set_fanout(border_cap, interior_cap); });
  return GMT_OK;
}

int gmt_set_tile_size(gmt_index* index, uint32_t tile_size) {
  if (!index) return bad("gmt_set_tile_size: null handle");
  if (!valid_tile(tile_size)) return bad("gmt_set_tile_size: tile_size must be 16, 32 or 64");
  B(index)->tile = tile_size;
  return GMT_OK;
}

uint32_t gmt_tile_size(gmt_index* index) { return index ? B(index)->tile : 0; }

struct gmt_server {
  std::unique_ptr<indexingpu_amd::masstree_server<tree_t<16>>> s16;
  std::unique_ptr<indexingpu_amd::masstree_server<tree_t<32>>> s32;
  std::unique_ptr<indexingpu_amd::masstree_server<tree_t<64>>> s64;
};

int gmt_server_create(gmt_index* index, const gmt_serve_config* config, gmt_server** out) {
  if (!index || !config || !out) return bad("gmt_server_create: null argument");
  *out = nullptr;
  if (config->max_wave_ops == 0 || config->ring_rows == 0 || config->epoch_waves == 0)
    return bad("gmt_server_create: max_wave_ops, ring_rows and epoch_waves must be > 0");
  indexingpu_amd::serve_config c;
  c.streams = config->streams;
  c.slabs_per_write = config->slabs_per_write;
  c.max_wave_ops = config->max_wave_ops;
  c.max_wave_scans = config->max_wave_scans;
  c.ring_rows = config->ring_rows;
  c.epoch_waves = config->epoch_waves;
  c.coalesce = config->coalesce;
  c.ordered = config->ordered;
  c.plain_batch = config->plain_batch;
  gmt_server* sv = new gmt_server();
  gmt_impl* p = B(index);
  int rc = guarded([&] {
    switch (p->tile) {
      case 32: sv->s32.reset(new indexingpu_amd::masstree_server<tree_t<32>>(*p->t32, c)); break;
      case 64: sv->s64.reset(new indexingpu_amd::masstree_server<tree_t<64>>(*p->t64, c)); break;
      default: sv->s16.reset(new indexingpu_amd::masstree_server<tree_t<16>>(*p->t16, c)); break;
    }
  });
  if (rc != GMT_OK) {
    delete sv;
    return rc;
  }
  *out = sv;
  return GMT_OK;
}

int gmt_server_destroy(gmt_server* server) {
  if (!server) return bad("gmt_server_destroy: null handle");
  return guarded([&] { delete server; });
}

int gmt_server_run(gmt_server* server, const gmt_serve_arena* arena,
                   const gmt_serve_request* requests, uint64_t n, gmt_serve_outcome* outcomes,
                   gmt_serve_stats* stats) {
  if (!server || !arena || (n && (!requests || !outcomes)))
    return bad("gmt_server_run: null argument");
  if (n && (!arena->types || !arena->keys || !arena->values || !arena->value_lengths ||
            !arena->results))
    return bad("gmt_server_run: the arena needs types, keys, values, value_lengths, results");
  indexingpu_amd::serve_arena a;
  a.types = arena->types;
  a.keys = arena->keys;
  a.max_key_length = arena->max_key_length;
  a.key_lengths = arena->key_lengths;
  a.values = arena->values;
  a.max_value_length = arena->max_value_length;
  a.value_lengths = arena->value_lengths;
  a.results = arena->results;
  a.scan_lower = arena->scan_lower;
  a.scan_lower_lengths = arena->scan_lower_lengths;
  a.scan_upper = arena->scan_upper;
  a.scan_upper_lengths = arena->scan_upper_lengths;
  a.scan_limits = arena->scan_limits;
  a.scan_max_count = arena->scan_max_count;
  a.scan_counts = arena->scan_counts;
  a.ring_keys = arena->ring_keys;
  a.ring_key_lengths = arena->ring_key_lengths;
  a.ring_values = arena->ring_values;
  a.ring_value_lengths = arena->ring_value_lengths;
  std::vector<indexingpu_amd::serve_request> rq(n);
  for (uint64_t i = 0; i < n; i++) {
    const gmt_serve_request& r = requests[i];
    if (r.num_scans && (!a.scan_lower || !a.scan_counts || !a.ring_keys || !a.ring_key_lengths ||
                        !a.ring_values || !a.ring_value_lengths))
      return bad("gmt_server_run: scans need scan_lower, scan_counts and the ring buffers");
    rq[i] = indexingpu_amd::serve_request{r.arrival_us, r.op_offset, r.scan_offset, r.num_ops,
                                          r.num_scans, r.num_writes};
  }
  std::vector<indexingpu_amd::serve_outcome> oc(n);
  indexingpu_amd::serve_stats st;
  int rc = guarded([&] {
    if (server->s16) st = server->s16->run(a, rq.data(), n, oc.data());
    if (server->s32) st = server->s32->run(a, rq.data(), n, oc.data());
    if (server->s64) st = server->s64->run(a, rq.data(), n, oc.data());
  });
  if (rc != GMT_OK) return rc;
  for (uint64_t i = 0; i < n; i++)
    outcomes[i] = gmt_serve_outcome{oc[i].admit_us, oc[i].done_us, oc[i].ring_row, oc[i].wave};
  if (stats) {
    stats->waves = st.waves;
    stats->ops = st.ops;
    stats->scans = st.scans;
    stats->flips = st.flips;
    stats->drains = st.drains;
    stats->census = st.census;
    stats->admission_waits = st.admission_waits;
    stats->elapsed_us = st.elapsed_us;
    stats->max_inflight = st.max_inflight;
    stats->device_error = st.device_error;
  }
  return GMT_OK;
}

const char* gmt_last_error_message(void) { return g_last_error.c_str(); }

}  // extern "C"
