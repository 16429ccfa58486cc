// GpuMasstree::gpu_masstree for MI355X (gfx950): the reference's drop-in API
// (include/gpu_masstree.hpp:43-191 host calls, :194-971 device calls) over this
// project's node layout, state machine and kernels.
//
// Host calls are asynchronous on `stream` and take device (or pinned host) buffers,
// exactly as the reference's.  Extensions, all defaulted so existing call sites
// compile unchanged:
//   * insert/update/erase accept `bool* results` after `stream` (per-op success);
//   * scan accepts a per-query limit array after `stream`;
//   * tile_size may be 16, 32 or 64;
//   * last_error(), drain(), validate_tree(), dump_entries() for tests and tools.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_context.hpp"
#include "hip_util.hpp"
#include "host_walk.hpp"
#include "kernels.hpp"
#include "layout.hpp"
#include "narrow_find.hpp"
#include "reclaim.hpp"
#include "slab_alloc.hpp"
#include "tile_policy.hpp"
#include "tree_ops.hpp"

namespace kernels {
// include/macros.hpp:51-58 (values are part of the drop-in ABI)
enum request_type : uint8_t {
  request_type_insert = 0,
  request_type_erase = 1,
  request_type_find = 2,
  request_type_update = 3
};
}  // namespace kernels

namespace indexingpu_amd {

inline u32 device_cu_count() {
  int dev = 0, cus = 0;
  IGPU_HIP_TRY(hipGetDevice(&dev));
  IGPU_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return (u32)cus;
}

// Resident-block capacity of a kernel at kBlockThreads, computed once per kernel.
template <class K>
inline u32 resident_blocks(K kernel) {
  static std::mutex mu;
  static std::unordered_map<const void*, u32> cache;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(reinterpret_cast<const void*>(kernel));
  if (it != cache.end()) return it->second;
  int per_cu = 0;
  IGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlockThreads, 0));
  if (per_cu < 1) per_cu = 1;
  u32 v = (u32)per_cu * device_cu_count();
  cache[reinterpret_cast<const void*>(kernel)] = v;
  return v;
}

}  // namespace indexingpu_amd

namespace GpuMasstree {

using ::device_allocator_context;
using ::device_reclaimer_context;

template <typename Allocator, typename Reclaimer, uint32_t tile_size = 16>
struct gpu_masstree {
  using size_type = uint32_t;
  using key_slice_type = uint32_t;
  using value_slice_type = uint32_t;
  static constexpr uint32_t tile_size_ = tile_size;
  static_assert(tile_size == 16 || tile_size == 32 || tile_size == 64,
                "tile_size must be 16, 32 or 64 lanes");
  static constexpr auto branching_factor = 16;
  static constexpr size_type invalid_pointer = 0xFFFFFFFFu;

  using host_allocator_type = Allocator;
  using device_allocator_instance_type = typename Allocator::device_instance_type;
  using device_allocator_context_type = device_allocator_context<Allocator>;
  using host_reclaimer_type = Reclaimer;
  using device_reclaimer_instance_type = typename Reclaimer::device_instance_type;
  using device_reclaimer_context_type = device_reclaimer_context<Reclaimer>;

  gpu_masstree() = delete;
  gpu_masstree(const host_allocator_type& host_allocator, const host_reclaimer_type& host_reclaimer)
      : allocator_(host_allocator.get_device_instance()),
        reclaimer_(host_reclaimer.get_device_instance()),
        inserted_hint_(new uint64_t(0)) {
    allocate();
  }
  // shallow copy (kernels receive the tree by value, as in the reference)
  gpu_masstree(const gpu_masstree& other)
      : root_index_(other.root_index_),
        allocator_(other.allocator_),
        reclaimer_(other.reclaimer_),
        err_(other.err_),
        barrier_(other.barrier_),
        inserted_hint_(other.inserted_hint_),
        owner_(false),
        fanout_flags_(other.fanout_flags_),
        shortcut_on_(false) {}
  template <typename, typename, uint32_t>
  friend struct gpu_masstree;
  // Attaches to the tree of an index with another tile width: the node layout does not
  // depend on the tile width, so both operate on the same tree (the owner frees it).
  template <uint32_t other_tile>
  explicit gpu_masstree(const gpu_masstree<Allocator, Reclaimer, other_tile>& other)
      : root_index_(other.root_index_),
        allocator_(other.allocator_),
        reclaimer_(other.reclaimer_),
        err_(other.err_),
        barrier_(other.barrier_),
        inserted_hint_(other.inserted_hint_),
        owner_(false),
        fanout_flags_(other.fanout_flags_),
        shortcut_on_(false) {}
  gpu_masstree& operator=(const gpu_masstree&) = delete;
  ~gpu_masstree() {
    if (scan_stack_) (void)hipFree(scan_stack_);  // per instance, never shared
    if (park_) (void)hipFree(park_);
    if (hints_) (void)hipFree(hints_);
    if (shortcut_) (void)hipFree(shortcut_);
    if (owner_) {
      if (err_) (void)hipFree(err_);
      if (barrier_) (void)hipFree(barrier_);
      delete inserted_hint_;
    }
  }

  // ------------------------------------------------------------------ host API
  // if key_lengths == nullptr, every key has max_key_length slices
  template <bool concurrent = false, bool use_shmem_key = false>
  void find(const key_slice_type* keys, const size_type max_key_length,
            const size_type* key_lengths, value_slice_type* values,
            const size_type max_value_length, size_type* value_lengths,
            const size_type num_keys, hipStream_t stream = 0) {
    if (!key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_keys);
    a.kind = indexingpu_amd::op_find;
    a.max_value_length = max_value_length;
    a.out_values = values;
    a.out_value_lengths = value_lengths;
    if (concurrent) {
      launch_batch<true>(a, stream, false);
    } else {
      refresh_shortcut(stream);
      switch (find_tile_) {
        case 1: launch_narrow<1>(a, stream); break;
        case 2: launch_narrow<2>(a, stream); break;
        case 4: launch_narrow<4>(a, stream); break;
        case 8: launch_narrow<8>(a, stream); break;
        case 16: launch_narrow<16>(a, stream); break;
        default: launch_wide(a, stream); break;
      }
    }
  }

  // find<false> whose two-pass scratch (one word per key) is the caller's `scratch` of
  // `scratch_words` words instead of the index's own buffer: nothing is allocated, so
  // it may run where allocation is not possible (graph capture, serving loops).  The
  // caller keeps num_keys <= scratch_words; calls sharing one scratch must not overlap.
  void find_with_scratch(const key_slice_type* keys, const size_type max_key_length,
                         const size_type* key_lengths, value_slice_type* values,
                         const size_type max_value_length, size_type* value_lengths,
                         const size_type num_keys, uint32_t* scratch, uint64_t scratch_words,
                         hipStream_t stream = 0) {
    if (num_keys > scratch_words || !key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_keys);
    a.kind = indexingpu_amd::op_find;
    a.max_value_length = max_value_length;
    a.out_values = values;
    a.out_value_lengths = value_lengths;
    switch (find_tile_) {
      case 1: launch_narrow<1>(a, stream, false, 0, 0, scratch); break;
      case 4: launch_narrow<4>(a, stream, false, 0, 0, scratch); break;
      case 8: launch_narrow<8>(a, stream, false, 0, 0, scratch); break;
      case 16: launch_narrow<16>(a, stream, false, 0, 0, scratch); break;
      case 2: launch_narrow<2>(a, stream, false, 0, 0, scratch); break;
      default: launch_wide(a, stream); break;
    }
  }
  // Fills the per-kernel occupancy cache of every read-only find kernel, so that a
  // later find_with_scratch makes no host allocation either.
  void warm_find_launches() {
    using namespace indexingpu_amd;
    auto warm_t = [](auto tag) {
      constexpr uint32_t T = decltype(tag)::value;
      resident_blocks(narrow_descend_kernel<T, 1>);
      resident_blocks(narrow_descend_kernel<T, 2>);
      resident_blocks(narrow_descend_kernel<T, 4>);
      resident_blocks(narrow_record_kernel<T>);
      resident_blocks(narrow_find_kernel<T, 1>);
      resident_blocks(narrow_find_kernel<T, 2>);
      resident_blocks(narrow_find_kernel<T, 4>);
      resident_blocks(narrow_find_kernel<T, 8>);
    };
    warm_t(std::integral_constant<uint32_t, 1>{});
    warm_t(std::integral_constant<uint32_t, 2>{});
    warm_t(std::integral_constant<uint32_t, 4>{});
    warm_t(std::integral_constant<uint32_t, 8>{});
    warm_t(std::integral_constant<uint32_t, 16>{});
    resident_blocks(find_kernel<tile_size, 1>);
    resident_blocks(find_kernel<tile_size, 2>);
    resident_blocks(find_kernel<tile_size, 4>);
    resident_blocks(find_kernel<tile_size, 8>);
  }

  // Lookups kept in flight per tile by the read-only find kernels (1, 2, 4 or 8).
  void set_find_depth(uint32_t k) { find_depth_ = (k == 1 || k == 2 || k == 8) ? k : 4; }
  uint32_t find_depth() const { return find_depth_; }
  // Lanes per lookup of read-only finds: 4, 8 or 16 (narrow tiles, each lane holding
  // 16 / width slots of a node), or 0 for the index's own tile width.
  void set_find_tile(uint32_t t) {
    find_tile_ = (t == 1 || t == 2 || t == 4 || t == 8 || t == 16) ? t : 0;
  }
  uint32_t find_tile() const { return find_tile_; }
  // Narrow finds in one kernel (1) or as a descent pass and a record pass (2).
  void set_find_passes(uint32_t p) { find_passes_ = p == 1 ? 1u : 2u; }
  uint32_t find_passes() const { return find_passes_; }
  // Lanes per lookup of the record pass of two-pass finds: 1, 2, 4, 8 or 16, or 0 for
  // the find tile.
  void set_record_tile(uint32_t t) {
    record_tile_ = (t == 1 || t == 2 || t == 4 || t == 8 || t == 16) ? t : 0;
  }
  uint32_t record_tile() const { return record_tile_; }

  template <bool update_if_exists = false, bool enable_suffix = true, bool use_shmem_key = false>
  void insert(const key_slice_type* keys, const size_type max_key_length,
              const size_type* key_lengths, const value_slice_type* values,
              const size_type max_value_length, const size_type* value_lengths,
              const size_type num_keys, hipStream_t stream = 0, bool* results = nullptr) {
    if (!key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_keys);
    a.kind = update_if_exists ? indexingpu_amd::op_upsert : indexingpu_amd::op_insert;
    a.flags = enable_suffix ? 0u : indexingpu_amd::kReqNoSuffix;
    a.values = values;
    a.max_value_length = max_value_length;
    a.value_lengths = value_lengths;
    a.results = results;
    launch_growing(a, stream);
  }

  template <bool enable_suffix = true, bool use_shmem_key = false>
  void update(const key_slice_type* keys, const size_type max_key_length,
              const size_type* key_lengths, const value_slice_type* values,
              const size_type max_value_length, const size_type* value_lengths,
              const size_type num_keys, hipStream_t stream = 0, bool* results = nullptr) {
    if (!key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_keys);
    a.kind = indexingpu_amd::op_update;
    a.values = values;
    a.max_value_length = max_value_length;
    a.value_lengths = value_lengths;
    a.results = results;
    launch_hinted(a, stream);
  }

  // Erase folds a border node it leaves empty into a neighbour (do_merge) and removes
  // layer roots it leaves empty (do_remove_empty_root), as the reference's switches ask;
  // pessimistic_merge (when the reference merges: on the way down or on a retry) has no
  // counterpart here, since nodes are merged once they are empty.  Results are the same
  // under every setting.
  template <bool do_merge, bool do_remove_empty_root>
  static constexpr uint32_t erase_flags() {
    return (do_merge ? 0u : indexingpu_amd::kReqNoMerge) |
           (do_remove_empty_root ? 0u : indexingpu_amd::kReqKeepEmptyLayers);
  }

  template <bool do_remove_empty_root = true, bool pessimistic_merge = true, bool do_merge = true,
            bool concurrent = true, bool use_shmem_key = false>
  void erase(const key_slice_type* keys, const size_type max_key_length,
             const size_type* key_lengths, const size_type num_keys, hipStream_t stream = 0,
             bool* results = nullptr) {
    if (!key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_keys);
    a.kind = indexingpu_amd::op_erase;
    a.flags = erase_flags<do_merge, do_remove_empty_root>();
    a.results = results;
    launch_hinted(a, stream);
  }

  template <bool use_upper_key = true, bool concurrent = false, bool use_shmem_key = false>
  void scan(const key_slice_type* lower_keys, const size_type* lower_key_lengths,
            const size_type max_key_length, const size_type max_count_per_query,
            const size_type num_queries, const key_slice_type* upper_keys = nullptr,
            const size_type* upper_key_lengths = nullptr, size_type* counts = nullptr,
            value_slice_type* values = nullptr, const size_type max_value_length = 1,
            size_type* value_lengths = nullptr, key_slice_type* out_keys = nullptr,
            size_type* out_key_lengths = nullptr, hipStream_t stream = 0,
            const size_type* per_query_limits = nullptr) {
    if (num_queries == 0 || !key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::scan_args a{};
    a.lower_keys = lower_keys;
    a.lower_key_lengths = lower_key_lengths;
    a.max_key_length = max_key_length;
    a.max_count = max_count_per_query;
    a.num_queries = num_queries;
    a.upper_keys = use_upper_key ? upper_keys : nullptr;
    a.upper_key_lengths = use_upper_key ? upper_key_lengths : nullptr;
    a.counts = counts;
    a.values = values;
    a.max_value_length = max_value_length;
    a.value_lengths = value_lengths;
    a.out_keys = out_keys;
    a.out_key_lengths = out_key_lengths;
    a.limits = per_query_limits;
    auto kernel = indexingpu_amd::scan_kernel<tile_size, concurrent>;
    const uint32_t tiles_per_block = indexingpu_amd::kBlockThreads / tile_size;
    uint64_t want = (num_queries + tiles_per_block - 1) / tiles_per_block;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(kernel));
    a.stack_layers = std::max<uint32_t>(max_key_length, 64u) + 1;
    ensure_scan_stack((uint64_t)blocks * tiles_per_block * a.stack_layers * 3u);
    a.stack = scan_stack_;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                       view(), a);
  }

  template <bool enable_suffix = true, bool erase_do_remove_empty_root = true,
            bool erase_pessimistic_merge = true, bool erase_do_merge = true,
            bool use_shmem_key = false>
  void mixed_batch(const kernels::request_type* request_types, const key_slice_type* keys,
                   const size_type max_key_length, const size_type* key_lengths,
                   value_slice_type* values, const size_type max_value_length,
                   size_type* value_lengths, bool* results, const size_type num_requests,
                   hipStream_t stream = 0) {
    if (!key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::batch_args a = base_args(keys, max_key_length, key_lengths, num_requests);
    a.types = reinterpret_cast<const uint8_t*>(request_types);
    a.flags = (enable_suffix ? 0u : indexingpu_amd::kReqNoSuffix) |
              erase_flags<erase_do_merge, erase_do_remove_empty_root>();
    a.kind = indexingpu_amd::op_find;
    a.values = values;
    a.max_value_length = max_value_length;
    a.value_lengths = value_lengths;
    a.out_values = values;
    a.out_value_lengths = value_lengths;
    a.results = results;
    if (num_requests >= mixed_split_min_) {
      // the finds as a read-only pass (narrow two-pass kernels, which leave every other
      // request alone), then the other requests; stream order separates the passes.  A
      // batch writes no key twice and reads no key it writes, so the outcome equals
      // that of one launch running all requests together.
      ensure_hints(num_requests);
      a.hints = hints_;
      refresh_shortcut(stream);
      launch_finds_pass(a, stream);
      a.skip_finds = 1;
    }
    launch_batch<true>(a, stream, true);
  }

  // Allocates the two-pass find scratch for batches of up to max_requests requests now
  // (it otherwise grows at the first larger batch, which synchronizes; callers that
  // capture launches into a graph reserve first).
  void reserve_find_scratch(uint64_t max_requests) {
    ensure_park(max_requests);
    ensure_hints(max_requests);
  }

  // Node fanout (an extension; the reference's nodes are fixed at 16 slots): later writes
  // split a border node once it holds border_cap entries (2..14) and an interior node once
  // it has interior_cap children (3..15); 0 keeps the layout's maximum.  The nodes of one
  // tree may have been built under different settings: readers use each node's own count.
  // Returns false (nothing changed) for a value out of range.
  bool set_fanout(uint32_t border_cap, uint32_t interior_cap) {
    if ((border_cap != 0 && (border_cap < 2 || border_cap > indexingpu_amd::kBorderCap)) ||
        (interior_cap != 0 && (interior_cap < 3 || interior_cap > indexingpu_amd::kInteriorCap)))
      return false;
    fanout_flags_ = (border_cap << indexingpu_amd::kReqBorderCapShift) |
                    (interior_cap << indexingpu_amd::kReqInteriorCapShift);
    return true;
  }
  uint32_t border_fanout() const {
    const uint32_t c = (fanout_flags_ >> indexingpu_amd::kReqBorderCapShift) & 15u;
    return c ? c : indexingpu_amd::kBorderCap;
  }
  uint32_t interior_fanout() const {
    const uint32_t c = (fanout_flags_ >> indexingpu_amd::kReqInteriorCapShift) & 15u;
    return c ? c : indexingpu_amd::kInteriorCap;
  }
  uint32_t fanout_flags() const { return fanout_flags_; }

  // Read-only descents of large trees (find, the read-only pass of mixed batches, the
  // start hints of write batches) start at a layer-0 shortcut table (refresh_shortcut)
  // instead of the root: on (the default) or off (the table is dropped).  An extension;
  // results do not depend on it.
  // Only the instance that owns the tree builds shortcut tables: merges rewrite the
  // border-level table of the launch that makes them, so a second table (of an instance
  // attached with another tile width) would miss the merges of the first one's writes.
  // The attached instances' writes give no table and count their merges, which retires
  // the owner's border-level table until its next rebuild.
  void set_find_shortcut(bool on) {
    shortcut_on_ = on && owner_;
    if (!on && shortcut_) {
      IGPU_HIP_TRY(hipDeviceSynchronize());
      IGPU_HIP_TRY(hipFree(shortcut_));
      shortcut_ = nullptr;
      shortcut_bits_ = shortcut_capacity_bits_ = 0;
      shortcut_size_ = 0;
    }
  }
  bool find_shortcut() const { return shortcut_on_; }
  uint32_t find_shortcut_bits() const { return shortcut_ ? shortcut_bits_ : 0u; }

  // Insert, update and erase batches of at least this many requests start their writes
  // at border nodes found by a read-only pass (UINT32_MAX: never).
  void set_write_hints(uint32_t min_requests) { write_hint_min_ = min_requests; }
  uint32_t write_hints() const { return write_hint_min_; }

  // Mixed batches of at least this many requests run their finds as a separate
  // read-only pass (0: every batch; UINT32_MAX: none).
  void set_mixed_split(uint32_t min_requests) { mixed_split_min_ = min_requests; }
  uint32_t mixed_split() const { return mixed_split_min_; }

  // Batched form (no counterpart in the reference): applies `num_batches` mixed batches
  // in order within one persistent launch.  The arrays hold the requests of every batch
  // back to back, with mixed_batch's layout; `offsets` is a device array of
  // num_batches + 1 u64 request offsets.  Same per-request outputs as mixed_batch.
  void multi_batch(const kernels::request_type* request_types, const key_slice_type* keys,
                   const size_type max_key_length, const size_type* key_lengths,
                   value_slice_type* values, const size_type max_value_length,
                   size_type* value_lengths, bool* results, const uint64_t* offsets,
                   const size_type num_batches, const uint64_t total_requests,
                   hipStream_t stream = 0) {
    if (num_batches == 0 || !key_length_ok(max_key_length, stream)) return;
    indexingpu_amd::multi_args m{};
    m.a = base_args(keys, max_key_length, key_lengths, total_requests);
    m.a.types = reinterpret_cast<const uint8_t*>(request_types);
    m.a.kind = indexingpu_amd::op_find;
    m.a.values = values;
    m.a.max_value_length = max_value_length;
    m.a.value_lengths = value_lengths;
    m.a.out_values = values;
    m.a.out_value_lengths = value_lengths;
    m.a.results = results;
    m.a.flags |= fanout_flags_;
    if (multi_mode_ == 1 && total_requests < (1ull << 32)) {
      refresh_shortcut(stream);
      // one batch-kernel launch per batch, each reading its range from `offsets` on the
      // device; stream order separates the batches, and the records each batch retires
      // are drained right after it
      m.a.results = results;
      auto kernel = indexingpu_amd::batch_kernel<tile_size, true>;
      const uint32_t blocks = std::max<uint32_t>(
          1, (uint32_t)((uint64_t)indexingpu_amd::resident_blocks(kernel) * multi_grid_percent_ / 100));
      // Batches of mixed_split_min_ requests or more run their finds as a read-only
      // pass first, as mixed_batch does.  Outside graph capture the offsets are read
      // back once, so each batch's grids fit its size and the split is decided per
      // batch; under capture the launches take full grids and the split follows the
      // average batch size.
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      IGPU_HIP_TRY(hipStreamIsCapturing(stream, &cs));
      std::vector<uint64_t> off;
      if (cs == hipStreamCaptureStatusNone) {
        off.resize((size_t)num_batches + 1);
        IGPU_HIP_TRY(hipMemcpyAsync(off.data(), offsets, off.size() * 8, hipMemcpyDeviceToHost,
                                    stream));
        IGPU_HIP_TRY(hipStreamSynchronize(stream));
      }
      const bool split_all = mixed_split_min_ != 0xFFFFFFFFu &&
                             total_requests >= (uint64_t)mixed_split_min_ * num_batches;
      // the find scratch is indexed from each batch's first request
      uint64_t largest = total_requests;
      if (!off.empty()) {
        largest = 0;
        for (uint32_t b = 0; b < num_batches; b++)
          largest = std::max<uint64_t>(largest, off[b + 1] - std::min(off[b], off[b + 1]));
        largest = std::min<uint64_t>(largest, total_requests);
      }
      if (mixed_split_min_ != 0xFFFFFFFFu) {
        ensure_park(largest);
        ensure_hints(largest);
      }
      const uint32_t tiles_per_block = indexingpu_amd::kBlockThreads / tile_size;
      for (uint32_t b = 0; b < num_batches; b++) {
        indexingpu_amd::batch_args a = m.a;
        a.range = offsets + b;
        uint64_t nb = total_requests;  // requests of this batch (an upper bound under capture)
        bool split = split_all;
        if (!off.empty()) {
          const uint64_t hi = std::min<uint64_t>(off[b + 1], total_requests);
          nb = hi - std::min<uint64_t>(off[b], hi);
          if (nb == 0) continue;
          split = mixed_split_min_ != 0xFFFFFFFFu && nb >= mixed_split_min_;
        }
        if (split) {
          a.hints = hints_;
          launch_finds_pass(a, stream, nb);
          a.skip_finds = 1;
        }
        const uint32_t bb =
            (uint32_t)std::min<uint64_t>(blocks, (nb + tiles_per_block - 1) / tiles_per_block);
        hipLaunchKernelGGL(kernel, dim3(bb), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                           view(), a);
        // slabs of the records this batch retired go back to the allocator before the
        // next batch's inserts take slabs (as with one mixed_batch call per batch)
        drain(stream);
      }
      return;
    }
    m.offsets = offsets;
    m.num_batches = num_batches;
    m.barrier = barrier_;
    auto kernel = indexingpu_amd::multi_batch_kernel<tile_size>;
    // every block must be resident at once (the kernel has a grid-wide barrier)
    const uint32_t blocks = std::max<uint32_t>(
        1, (uint32_t)((uint64_t)indexingpu_amd::resident_blocks(kernel) * multi_grid_percent_ / 100));
    m.groups = std::min<uint32_t>(blocks, indexingpu_amd::kMaxBarrierGroups);
    hipLaunchKernelGGL(indexingpu_amd::zero_kernel, dim3(1), dim3(256), 0, stream, m.barrier,
                       indexingpu_amd::kBarrierWords);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                       view(), m);
    drain(stream);
  }
  // Execution of the batched form: 0 = one persistent launch with a grid barrier between
  // batches, 1 = one launch per batch (no barrier, one drain at the end).
  void set_multi_mode(uint32_t mode) { multi_mode_ = mode ? 1u : 0u; }
  // Fraction (percent) of the resident-block capacity the batched form launches.
  void set_multi_grid_percent(uint32_t p) { multi_grid_percent_ = std::min<uint32_t>(std::max<uint32_t>(p, 1), 100); }

  // ---------------------------------------------------------------- device API
  // Every lane of the tile calls these with the same arguments.  Keys and values are
  // plain slice pointers.  Each call runs the same step machine as the batch kernels.
  template <bool concurrent, typename tile_type, typename K, typename V>
  __device__ size_type cooperative_find(K& key, size_type key_length, V& value,
                                        size_type max_value_length, const tile_type& tile,
                                        device_allocator_context_type& allocator) {
    indexingpu_amd::wave_tile<tile_size, concurrent> c(allocator.view, indexingpu_amd::retire_view{},
                                                      err_);
    indexingpu_amd::request rq{indexingpu_amd::op_find, key_length, as_ptr(key), nullptr, 0,
                               as_mut_ptr(value), max_value_length};
    return run_op<concurrent>(c, rq);
  }

  // Range query from a user kernel (the reference's cooperative_scan,
  // include/gpu_masstree.hpp:250-395): up to out_max_count entries with key >= the lower
  // key and, with use_upper_key, <= the upper key, ascending; entry c's value goes to
  // out_values + c * out_value_max_length (min(len, max) slices) with its full length in
  // out_value_lengths[c], its key likewise to out_keys / out_key_lengths; any output may
  // be null.  Returns the number of entries.  The layer stack of keys that span trie
  // layers takes slabs from the tile's allocator context while the scan runs.
  template <bool use_upper_key, bool concurrent, typename tile_type, typename K>
  __device__ size_type cooperative_scan(K& lower_key, const size_type lower_key_length,
                                        const tile_type& tile,
                                        device_allocator_context_type& allocator,
                                        const key_slice_type* upper_key = nullptr,
                                        const size_type upper_key_length = 1,
                                        size_type out_max_count = 1,
                                        value_slice_type* out_values = nullptr,
                                        size_type* out_value_lengths = nullptr,
                                        const size_type out_value_max_length = 1,
                                        key_slice_type* out_keys = nullptr,
                                        size_type* out_key_lengths = nullptr,
                                        const size_type out_key_max_length = 1) {
    indexingpu_amd::wave_tile<tile_size, concurrent> c(allocator.view,
                                                       indexingpu_amd::retire_view{}, err_);
    if (out_max_count == 0) return 0;
    if (lower_key_length == 0 || (use_upper_key && (upper_key == nullptr || upper_key_length == 0))) {
      c.set_error(indexingpu_amd::kErrBadLength);
      return 0;
    }
    c.ac = allocator.cache;
    indexingpu_amd::scan_query<tile_size> q;
    q.lo = as_ptr(lower_key);
    q.nlo = lower_key_length;
    q.use_hi = use_upper_key;
    q.hi = use_upper_key ? upper_key : nullptr;
    q.nhi = use_upper_key ? upper_key_length : 0u;
    q.limit = out_max_count;
    q.vals = out_values;
    q.vlens = out_value_lengths;
    q.keys = out_keys;
    q.klens = out_key_lengths;
    q.maxV = out_value_max_length;
    q.maxK = out_key_max_length;
    indexingpu_amd::slab_scan_stack stk;
    const size_type n = indexingpu_amd::tile_scan<tile_size, concurrent>(c, root_index_, q, stk);
    stk.release(c);
    allocator.cache = c.ac;
    return n;
  }

  template <bool update_if_exists = false, bool enable_suffix = true, typename tile_type,
            typename K, typename V>
  __device__ bool cooperative_insert(K& key, const size_type key_length, V& value,
                                     const size_type value_length, const tile_type& tile,
                                     device_allocator_context_type& allocator,
                                     device_reclaimer_context_type& reclaimer) {
    indexingpu_amd::wave_tile<tile_size, true> c(allocator.view, reclaimer.view, err_, root_index_);
    c.ac = allocator.cache;
    c.rc = reclaimer.cache;
    indexingpu_amd::request rq{update_if_exists ? indexingpu_amd::op_upsert : indexingpu_amd::op_insert,
                               key_length, as_ptr(key), as_ptr(value), value_length, nullptr, 0,
                               (enable_suffix ? 0u : indexingpu_amd::kReqNoSuffix) | fanout_flags_};
    const bool r = run_op<true>(c, rq) != 0;
    allocator.cache = c.ac;
    reclaimer.cache = c.rc;
    return r;
  }

  template <bool enable_suffix = true, typename tile_type, typename K, typename V>
  __device__ bool cooperative_update(K& key, const size_type key_length, V& value,
                                     const size_type value_length, const tile_type& tile,
                                     device_allocator_context_type& allocator,
                                     device_reclaimer_context_type& reclaimer) {
    indexingpu_amd::wave_tile<tile_size, true> c(allocator.view, reclaimer.view, err_, root_index_);
    c.ac = allocator.cache;
    c.rc = reclaimer.cache;
    indexingpu_amd::request rq{indexingpu_amd::op_update, key_length, as_ptr(key), as_ptr(value),
                               value_length, nullptr, 0, fanout_flags_};
    const bool r = run_op<true>(c, rq) != 0;
    allocator.cache = c.ac;
    reclaimer.cache = c.rc;
    return r;
  }

  template <bool concurrent, bool do_merge, bool pessimistic_merge, bool do_remove_empty_root,
            typename tile_type, typename K>
  __device__ bool cooperative_erase(K& key, const size_type key_length, const tile_type& tile,
                                    device_allocator_context_type& allocator,
                                    device_reclaimer_context_type& reclaimer) {
    indexingpu_amd::wave_tile<tile_size, true> c(allocator.view, reclaimer.view, err_, root_index_);
    c.ac = allocator.cache;
    c.rc = reclaimer.cache;
    indexingpu_amd::request rq{indexingpu_amd::op_erase, key_length, as_ptr(key), nullptr, 0,
                               nullptr, 0, erase_flags<do_merge, do_remove_empty_root>() | fanout_flags_};
    const bool r = run_op<true>(c, rq) != 0;
    allocator.cache = c.ac;
    reclaimer.cache = c.rc;
    return r;
  }

  // -------------------------------------------------------------- host extras
  // Waits for the device, returns the first error code since the last call (0 = none)
  // and clears it.
  uint32_t last_error() {
    IGPU_HIP_TRY(hipDeviceSynchronize());
    uint32_t h = 0;
    IGPU_HIP_TRY(hipMemcpy(&h, err_, 4, hipMemcpyDeviceToHost));
    clear_error();
    return h;
  }
  // Clears the error word and the livelock guard's trip record.
  void clear_error() {
    IGPU_HIP_TRY(hipMemset(err_, 0, 4));
    IGPU_HIP_TRY(hipMemset(err_ + indexingpu_amd::kTripOffset, 0, indexingpu_amd::kTripWords * 4));
  }
  // Waits for the device and copies the livelock guard's record of the operation it
  // abandoned (indexingpu_amd::kTripWords words, layout in layout.hpp) into out; false
  // when the guard has not fired since the last clear.
  bool last_guard_trip(uint32_t* out) {
    IGPU_HIP_TRY(hipDeviceSynchronize());
    IGPU_HIP_TRY(hipMemcpy(out, err_ + indexingpu_amd::kTripOffset, indexingpu_amd::kTripWords * 4,
                           hipMemcpyDeviceToHost));
    return out[0] == 1;
  }

  // Frees the records retired so far (batch calls do this themselves).
  void drain(hipStream_t stream = 0) {
    if (!reclaimer_.enabled) return;
    hipLaunchKernelGGL(indexingpu_amd::drain_kernel, dim3(256), dim3(256), 0, stream,
                       allocator_, reclaimer_);
    hipLaunchKernelGGL(indexingpu_amd::reset_kernel, dim3(1), dim3(64), 0, stream, reclaimer_);
  }

  // Walks the whole tree on the host (copy of the used part of the pool).
  indexingpu_amd::walk_stats walk(const indexingpu_amd::entry_emit_fn& emit) {
    IGPU_HIP_TRY(hipDeviceSynchronize());
    std::vector<uint32_t> bm(allocator_.num_words);
    IGPU_HIP_TRY(hipMemcpy(bm.data(), allocator_.bitmap, 4ull * bm.size(), hipMemcpyDeviceToHost));
    uint64_t top = 0;
    for (uint64_t w = 0; w < bm.size(); w++)
      if (bm[w]) top = w + 1;
    std::vector<uint64_t> img(top * 32 * indexingpu_amd::kSlots);
    if (top)
      IGPU_HIP_TRY(hipMemcpy(img.data(), allocator_.pool, img.size() * 8, hipMemcpyDeviceToHost));
    const uint64_t nslabs = top * 32;
    indexingpu_amd::host_walker w(
        [&](uint32_t idx) -> const uint64_t* {
          return idx < nslabs ? img.data() + (uint64_t)idx * indexingpu_amd::kSlots : nullptr;
        },
        emit);
    return w.run(root_index_);
  }
  bool validate_tree(std::string* msg = nullptr) {
    auto s = walk(nullptr);
    if (msg) *msg = s.error;
    return s.error.empty();
  }

  // reference debug API (gpu_masstree.hpp:1411-1472)
  void print() {
    walk([](const std::vector<uint32_t>& k, const std::vector<uint32_t>& v) {
      std::printf("key[");
      for (size_t i = 0; i < k.size(); i++) std::printf(i ? " %u" : "%u", k[i]);
      std::printf("] -> %zu value slices\n", v.size());
    });
  }
  void validate() {
    auto s = walk(nullptr);
    if (!s.error.empty()) std::printf("validate: %s\n", s.error.c_str());
    uint64_t total = (s.nodes + s.record_slabs) * indexingpu_amd::kSlabBytes;
    std::printf("%lu entries, %lu nodes (+%lu suffix nodes) found\n", (unsigned long)s.entries,
                (unsigned long)s.nodes, (unsigned long)s.record_slabs);
    std::printf("Total Space Consumption: %lu B (%f B/entry)\n", (unsigned long)total,
                s.entries ? (double)total / (double)s.entries : 0.0);
  }
  std::size_t num_nodes_used() { return allocator_.num_allocated_slabs(); }

  // The layer-0 shortcut table for read-only descents launched on `stream`, refreshed
  // first (which may synchronize the stream): {table, shift}, table null when there is
  // none.  For callers that launch the narrow descent kernels themselves (the served
  // form); a table stays correct under later writes.
  std::pair<const uint32_t*, uint32_t> find_shortcut_table(hipStream_t stream) {
    refresh_shortcut(stream);
    return {shortcut_, shortcut_ ? 32u - shortcut_bits_ : 0u};
  }
  // Its border-level twin and the merge count it was built at (tree_view::shortcut_border,
  // shortcut_epoch), null when there is none; same shift.  Only for descents that no
  // write of the index runs beside (a kernel checks the merge count once, at its start);
  // write launches given it rewrite its entries when they merge nodes away
  // (shortcut_patch).
  std::pair<uint32_t*, const uint32_t*> find_shortcut_border() const {
    if (!shortcut_) return {nullptr, nullptr};
    return {shortcut_ + (1ull << shortcut_capacity_bits_),
            shortcut_ + (2ull << shortcut_capacity_bits_)};
  }

  size_type root_index_ = 0;
  device_allocator_instance_type allocator_;
  device_reclaimer_instance_type reclaimer_;
  uint32_t* err_ = nullptr;
  uint32_t* barrier_ = nullptr;  // grid barrier area of the batched form

 private:
  template <class K>
  __host__ __device__ static const uint32_t* as_ptr(K& k) {
    return static_cast<const uint32_t*>(k);
  }
  template <class V>
  __host__ __device__ static uint32_t* as_mut_ptr(V& v) {
    return static_cast<uint32_t*>(v);
  }

  template <bool Validate, class Tile>
  __device__ size_type run_op(Tile& c,
                              const indexingpu_amd::request& rq) {
    if (rq.key_length == 0) {
      c.set_error(indexingpu_amd::kErrBadLength);
      return 0;
    }
    c.load_key_cache(rq);
    indexingpu_amd::op_state<indexingpu_amd::gpu_node> st;
    indexingpu_amd::op_begin(c, st, root_index_);
    uint32_t backoff = 0;
    while (st.phase != indexingpu_amd::ph_done) {
      if (indexingpu_amd::op_step<Validate>(c, rq, st) || indexingpu_amd::op_holds_locks(st)) {
        backoff = 0;
        continue;
      }
      if (backoff < indexingpu_amd::kMaxBackoffShift) backoff += 1;
      for (uint32_t k = 0; k < (1u << backoff); k++) __builtin_amdgcn_s_sleep(1);
      st.steps += 1u << backoff;
    }
    return st.result;
  }

  indexingpu_amd::tree_view view() const {
    indexingpu_amd::tree_view v{allocator_, reclaimer_, err_, root_index_, shortcut_,
                                shortcut_ ? 32u - shortcut_bits_ : 0u};
    const auto border = find_shortcut_border();
    v.shortcut_border = border.first;
    v.shortcut_epoch = border.second;
    return v;
  }

  // The layer-0 shortcut table of the read-only descents (tree_view::shortcut): built
  // once the tree holds kShortcutMinKeys keys (by the insert count) and its root is at
  // level 3 or above, with about one entry per key (2^18..2^26 entries, 4 B each), and
  // rebuilt when the count has doubled since.  Measured at 64M var-len keys (two-pass
  // finds, MI355X): a batch takes 9.1 ms from the root, 8.3 ms with one entry per 64 keys,
  // 7.3 with one per 4 and 6.9 with one per key (each prefix then starts at its own
  // level-1 node with no move-right).  A table left from before later writes is still
  // correct (see tree_view), only less precise.  Not built under graph capture (no
  // allocation there); an existing table is used.
  static constexpr uint64_t kShortcutMinKeys = 1ull << 18;
  static constexpr uint32_t kShortcutMaxBits = 26;
  void refresh_shortcut(hipStream_t stream) {
    if (!shortcut_on_) return;
    const uint64_t size = *inserted_hint_;
    if (size < kShortcutMinKeys || (shortcut_size_ && size < 2 * shortcut_size_)) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    IGPU_HIP_TRY(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return;
    shortcut_size_ = size;  // decided for this size, table or not
    // a root below level 3 leaves at most one interior level to skip
    uint64_t meta = 0;
    IGPU_HIP_TRY(hipStreamSynchronize(stream));
    IGPU_HIP_TRY(hipMemcpy(&meta, allocator_.pool + (uint64_t)root_index_ * indexingpu_amd::kSlots +
                                      indexingpu_amd::kMetaSlot, 8, hipMemcpyDeviceToHost));
    const uint32_t rh = indexingpu_amd::slot_lo(meta);
    if (indexingpu_amd::hdr_border(rh) || indexingpu_amd::hdr_level(rh) < 3) return;
    uint32_t bits = 18;
    while (bits < kShortcutMaxBits && (1ull << bits) < size) bits++;
    if (bits > shortcut_capacity_bits_) {
      // level-1 table, border-level table (2^capacity entries each), built-at merge count
      IGPU_HIP_TRY(hipDeviceSynchronize());
      if (shortcut_) IGPU_HIP_TRY(hipFree(shortcut_));
      shortcut_ = nullptr;
      IGPU_HIP_TRY(hipMalloc(&shortcut_, (8ull << bits) + 256));
      shortcut_capacity_bits_ = bits;
    }
    const uint32_t entries = 1u << bits;
    hipLaunchKernelGGL(indexingpu_amd::shortcut_build_kernel, dim3((entries + 255) / 256),
                       dim3(256), 0, stream, indexingpu_amd::tree_view{allocator_, reclaimer_,
                                                                       err_, root_index_},
                       shortcut_, shortcut_ + (1ull << shortcut_capacity_bits_),
                       shortcut_ + (2ull << shortcut_capacity_bits_), entries, 32u - bits);
    shortcut_bits_ = bits;
  }

  // Keys of more than 65,536 slices are refused (the device error word reports
  // kErrBadLength, the call does nothing): record lengths and trie layers are 16-bit.
  static constexpr uint32_t kMaxKeyLength = 65536;
  bool key_length_ok(uint32_t max_key_length, hipStream_t stream) {
    if (max_key_length <= kMaxKeyLength) return true;
    hipLaunchKernelGGL(indexingpu_amd::set_error_kernel, dim3(1), dim3(1), 0, stream, err_,
                       (uint32_t)indexingpu_amd::kErrBadLength);
    return false;
  }

  static indexingpu_amd::batch_args base_args(const uint32_t* keys, uint32_t maxK,
                                              const uint32_t* klens, uint64_t n) {
    indexingpu_amd::batch_args a{};
    a.keys = keys;
    a.max_key_length = maxK;
    a.key_lengths = klens;
    a.n = n;
    return a;
  }

  template <bool Validate>
  void launch_batch(const indexingpu_amd::batch_args& a0, hipStream_t stream, bool mutating) {
    if (a0.n == 0) return;
    indexingpu_amd::batch_args a = a0;
    a.flags |= fanout_flags_;
    auto kernel = indexingpu_amd::batch_kernel<tile_size, Validate>;
    const uint32_t tiles_per_block = indexingpu_amd::kBlockThreads / tile_size;
    uint64_t want = (a.n + tiles_per_block - 1) / tiles_per_block;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(kernel));
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                       view(), a);
    if (mutating) drain(stream);
  }

  // A large insert batch into a small tree would start every tile on the same few
  // nodes, where they serialize on the locks and on the coherent reads of those lines
  // (16M keys into an empty tree: 13 Mops/s in one launch, ~500 Mops/s in growing
  // launches).  The batch therefore runs as consecutive launches that each grow the
  // tree at most kGrowthFactor-fold (from kGrowthMin keys), then the rest in one launch.
  // The launches take the requests in a scattered order (index times a multiplier coprime to n, modulo n), so
  // each one is a sample of the whole batch: a batch in key order -- or one whose keys
  // all lie past the tree's current right edge -- then grows every key range at once
  // instead of sending every tile to the same edge node.  The requests of one batch are
  // independent, so any order gives the results of one launch.  The size estimate
  // counts every key inserted through this API (nodes are never merged, so it also
  // bounds how spread out the tree is).
  static constexpr uint64_t kGrowthMin = 512;
  static constexpr uint64_t kGrowthFactor = 4;
  // Weyl multiplier: the integer nearest n / golden ratio that is coprime to n, so
  // j -> j * m mod n is a bijection and any run of consecutive j lands spread evenly
  // over [0, n) (three-gap theorem).
  static uint64_t scatter_multiplier(uint64_t n) {
    if (n <= 2) return n == 2 ? 1 : 0;
    uint64_t m = (uint64_t)((double)n * 0.6180339887498949);
    if (m == 0) m = 1;
    auto gcd = [](uint64_t x, uint64_t y) {
      while (y) {
        const uint64_t t = x % y;
        x = y;
        y = t;
      }
      return x;
    };
    while (gcd(m, n) != 1) m++;
    return m;
  }
  void launch_growing(indexingpu_amd::batch_args a, hipStream_t stream) {
    uint64_t size = *inserted_hint_;
    const uint64_t n = a.n;
    if (n == 0) return;
    if (std::max<uint64_t>(kGrowthMin, size * (kGrowthFactor - 1)) >= n) {
      launch_hinted(a, stream);  // the tree is large enough: no growth stages
      *inserted_hint_ = size + n;
      return;
    }
    a.perm_n = n;
    a.perm_mul = scatter_multiplier(n);
    uint64_t done = 0;
    while (done < n) {
      uint64_t stage = std::max<uint64_t>(kGrowthMin, size * (kGrowthFactor - 1));
      if (stage > n - done) stage = n - done;
      indexingpu_amd::batch_args s = a;
      s.perm_lo = done;
      s.n = stage;
      *inserted_hint_ = size;  // the stage's hint descent may use a shortcut table
      launch_hinted(s, stream);
      done += stage;
      size += stage;
    }
    *inserted_hint_ = size;
  }

  // A write launch (insert, update or erase requests) in chunks of at most kHintChunk
  // positions.  With write hints on, each chunk first gets a read-only descent that
  // walks every request to its border node and leaves a start hint there, and the
  // chunk's writes then start at those nodes instead of the root (a B-link descent may
  // start at any node left of the key's on its level; nodes that split meanwhile are
  // left by moving right).  The requests of a batch are independent, so chunks give
  // the results of one launch.
  static constexpr uint64_t kHintChunk = 1ull << 22;
  void launch_hinted(const indexingpu_amd::batch_args& a, hipStream_t stream) {
    const uint64_t n = a.n;
    if (n == 0) return;
    if (write_hint_min_ == 0xFFFFFFFFu || n < write_hint_min_) {
      launch_batch<true>(a, stream, true);
      return;
    }
    ensure_hints(std::min<uint64_t>(n, kHintChunk));
    refresh_shortcut(stream);
    for (uint64_t off = 0; off < n; off += kHintChunk) {
      indexingpu_amd::batch_args c = a;
      c.n = std::min<uint64_t>(kHintChunk, n - off);
      if (a.perm_mul) {
        c.perm_lo = a.perm_lo + off;  // positions of a scattered order
      } else if (off) {
        c.keys = a.keys + off * a.max_key_length;
        if (a.key_lengths) c.key_lengths = a.key_lengths + off;
        if (a.values) c.values = a.values + off * a.max_value_length;
        if (a.value_lengths) c.value_lengths = a.value_lengths + off;
        if (a.results) c.results = a.results + off;
      }
      c.hints = hints_;
      switch (find_tile_) {
        case 1: launch_descend_only<1>(c, stream); break;
        case 4: launch_descend_only<4>(c, stream); break;
        case 8: launch_descend_only<8>(c, stream); break;
        case 16: launch_descend_only<16>(c, stream); break;
        default: launch_descend_only<2>(c, stream); break;
      }
      launch_batch<true>(c, stream, true);
    }
  }
  template <uint32_t T>
  void launch_descend_only(const indexingpu_amd::batch_args& a, hipStream_t stream) {
    constexpr uint32_t per_block = indexingpu_amd::kBlockThreads / T;
    // a scattered order (growth stages) reads random key rows either way: no ring
    auto kernel = a.perm_mul ? indexingpu_amd::narrow_descend_kernel<T, 1, true, false>
                             : indexingpu_amd::narrow_descend_kernel<T, 1, true>;
    const uint64_t want = (a.n + per_block - 1) / per_block;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(kernel));
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                       view(), a, park_);
  }

  template <typename Kern>
  void launch_reader(Kern kernel, uint32_t per_block, const indexingpu_amd::batch_args& a,
                     hipStream_t stream) {
    if (a.n == 0) return;
    uint64_t want = (a.n + per_block - 1) / per_block;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(kernel));
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                       view(), a);
  }
  // read-only pass over the finds of a mixed batch (or of the range a.range names, of
  // at most `work` requests: the grids are sized to it)
  // The read-only pass of a mixed batch (and of each batch of the batched form).  A batch
  // of up to kSmallFindsPass requests of short keys (at most kShortKey slices) holds a few
  // lookups per resident tile, so each tile keeps four in flight (depth 4) instead of
  // taking them one after another.  Measured on MI355X: 1M-request batches of 4-slice keys
  // over 64M keys, 2159-2166 Mops/s at depth 4 against 2012-2016 at depth 1.  Batches of
  // 16-slice keys under 12-slice shared prefixes (a walk through many trie layers) ran 7%
  // slower at depth 4, so they, and larger batches, use the tuned depth.  With a shortcut
  // table the chains are short enough that the tuned depth (1) wins there too: 2,480 Mops/s
  // against 2,412-2,434 at depth 4 on the same 1M-request batches.
  static constexpr uint64_t kSmallFindsPass = 1ull << 21;
  static constexpr uint32_t kShortKey = 4;
  void launch_finds_pass(const indexingpu_amd::batch_args& a, hipStream_t stream,
                         uint64_t work = 0) {
    const bool small = (work ? work : a.n) <= kSmallFindsPass && a.max_key_length <= kShortKey &&
                       !shortcut_;
    const uint32_t depth = small ? 4u : find_depth_;
    switch (find_tile_) {
      case 1: launch_narrow<1>(a, stream, true, work, depth); break;
      case 4: launch_narrow<4>(a, stream, true, work, depth); break;
      case 8: launch_narrow<8>(a, stream, true, work, depth); break;
      case 16: launch_narrow<16>(a, stream, true, work, depth); break;
      default: launch_narrow<2>(a, stream, true, work, depth); break;
    }
  }
  // Smallest batch whose general read-only pass takes positions from the wave's ring.  A
  // wave of 2-lane tiles holds one lookup per tile but takes 64-position chunks, so a batch
  // the resident grid covers in one round runs in two; on the 16-slice batched workload
  // (batches of 1K..1M requests) the ring on every batch measured 4% slower, while the 1M
  // mixed batches gained.
  static constexpr uint64_t kGeneralRingMin = 1ull << 19;
  template <uint32_t T>
  void launch_narrow(const indexingpu_amd::batch_args& a, hipStream_t stream,
                     bool two_pass = false, uint64_t work = 0, uint32_t depth = 0,
                     uint32_t* park = nullptr) {
    constexpr uint32_t tiles = indexingpu_amd::kBlockThreads / T;
    if (work == 0) work = a.n;
    if (depth == 0) depth = find_depth_;
    if ((two_pass || find_passes_ == 2) && a.n > 0) {
      if (!park) {
        ensure_park(a.range ? work : a.n);  // indexed from the first request of the range
        park = park_;
      }
      auto launch_descend = [&](auto kernel, uint32_t per_block) {
        const uint64_t want = (work + per_block - 1) / per_block;
        const uint32_t blocks =
            (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(kernel));
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream,
                           view(), a, park);
      };
      // plain find batches take the find-only kernels; the finds pass of a mixed batch
      // (kinds, ranges, write hints) the general ones
      const bool general = a.types || a.range || a.hints || a.perm_mul;
      if (general && (a.range || a.perm_mul || a.n < kGeneralRingMin)) {
        // the batched form's device-read ranges, scattered orders and batches below
        // kGeneralRingMin requests: tiles step through their own positions
        switch (depth) {
          case 1: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 1, true, false>, tiles); break;
          case 4: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 4, true, false>, tiles * 4); break;
          default: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 2, true, false>, tiles * 2); break;
        }
      } else if (general) {
        switch (depth) {
          case 1: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 1, true>, tiles); break;
          case 4: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 4, true>, tiles * 4); break;
          default: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 2, true>, tiles * 2); break;
        }
      } else {
        switch (depth) {
          case 1: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 1>, tiles); break;
          case 4: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 4>, tiles * 4); break;
          default: launch_descend(indexingpu_amd::narrow_descend_kernel<T, 2>, tiles * 2); break;
        }
      }
      switch (record_tile_ ? record_tile_ : T) {
        case 1: launch_record<1>(a, stream, work, park); break;
        case 2: launch_record<2>(a, stream, work, park); break;
        case 4: launch_record<4>(a, stream, work, park); break;
        case 8: launch_record<8>(a, stream, work, park); break;
        default: launch_record<16>(a, stream, work, park); break;
      }
      return;
    }
    switch (depth) {
      case 1: launch_reader(indexingpu_amd::narrow_find_kernel<T, 1>, tiles, a, stream); break;
      case 2: launch_reader(indexingpu_amd::narrow_find_kernel<T, 2>, tiles * 2, a, stream); break;
      case 8: launch_reader(indexingpu_amd::narrow_find_kernel<T, 8>, tiles * 8, a, stream); break;
      default: launch_reader(indexingpu_amd::narrow_find_kernel<T, 4>, tiles * 4, a, stream); break;
    }
  }
  template <uint32_t T>
  void launch_record(const indexingpu_amd::batch_args& a, hipStream_t stream, uint64_t work,
                     uint32_t* park) {
    constexpr uint32_t tiles = indexingpu_amd::kBlockThreads / T;
    auto rec = indexingpu_amd::narrow_record_kernel<T>;
    const uint64_t want = (work + tiles - 1) / tiles;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, indexingpu_amd::resident_blocks(rec));
    hipLaunchKernelGGL(rec, dim3(blocks), dim3(indexingpu_amd::kBlockThreads), 0, stream, view(),
                       a, static_cast<const uint32_t*>(park));
  }
  void launch_wide(const indexingpu_amd::batch_args& a, hipStream_t stream) {
    constexpr uint32_t tiles = indexingpu_amd::kBlockThreads / tile_size;
    switch (find_depth_) {
      case 1: launch_reader(indexingpu_amd::find_kernel<tile_size, 1>, tiles, a, stream); break;
      case 2: launch_reader(indexingpu_amd::find_kernel<tile_size, 2>, tiles * 2, a, stream); break;
      case 8: launch_reader(indexingpu_amd::find_kernel<tile_size, 8>, tiles * 8, a, stream); break;
      default: launch_reader(indexingpu_amd::find_kernel<tile_size, 4>, tiles * 4, a, stream); break;
    }
  }

  // per-lookup scratch of the two-pass find (grows on demand; allocating synchronizes)
  void ensure_park(uint64_t n) {
    if (n <= park_words_) return;
    IGPU_HIP_TRY(hipDeviceSynchronize());
    if (park_) IGPU_HIP_TRY(hipFree(park_));
    IGPU_HIP_TRY(hipMalloc(&park_, n * 4));
    park_words_ = n;
  }
  // start hints of the writes of split mixed batches (kHintWords words per request)
  void ensure_hints(uint64_t n) {
    if (n <= hint_requests_) return;
    IGPU_HIP_TRY(hipDeviceSynchronize());
    if (hints_) IGPU_HIP_TRY(hipFree(hints_));
    IGPU_HIP_TRY(hipMalloc(&hints_, n * 4 * indexingpu_amd::kHintWords));
    hint_requests_ = n;
  }

  void ensure_scan_stack(uint64_t words) {
    if (words <= scan_stack_words_) return;
    IGPU_HIP_TRY(hipDeviceSynchronize());
    if (scan_stack_) IGPU_HIP_TRY(hipFree(scan_stack_));
    IGPU_HIP_TRY(hipMalloc(&scan_stack_, words * 4));
    scan_stack_words_ = words;
  }

  void allocate() {
    IGPU_HIP_TRY(hipMalloc(&err_, 256));
    IGPU_HIP_TRY(hipMemset(err_, 0, 256));
    IGPU_HIP_TRY(hipMalloc(&barrier_, indexingpu_amd::kBarrierWords * 4));
    uint32_t* d_root = err_ + 1;
    hipLaunchKernelGGL(indexingpu_amd::init_root_kernel, dim3(1), dim3(64), 0, 0, allocator_, err_,
                       d_root);
    IGPU_HIP_TRY(hipGetLastError());
    IGPU_HIP_TRY(hipDeviceSynchronize());
    IGPU_HIP_TRY(hipMemcpy(&root_index_, d_root, 4, hipMemcpyDeviceToHost));
    if (root_index_ == invalid_pointer)
      throw indexingpu_amd::hip_error("gpu_masstree: slab pool exhausted at construction");
  }

  uint64_t* inserted_hint_ = nullptr;  // host counter shared with attached instances
  bool owner_ = true;
  uint32_t find_depth_ = 1;
  uint32_t find_tile_ = 2;
  uint32_t find_passes_ = 2;
  uint32_t record_tile_ = 0;
  uint32_t mixed_split_min_ = 1u << 14;
  uint32_t write_hint_min_ = 1;
  uint32_t* park_ = nullptr;
  uint64_t park_words_ = 0;
  uint32_t* hints_ = nullptr;
  uint64_t hint_requests_ = 0;
  uint32_t fanout_flags_ = 0;  // request flag bits of set_fanout (0: the layout's maximum)
  uint32_t* shortcut_ = nullptr;     // layer-0 shortcut table (refresh_shortcut)
  uint32_t shortcut_bits_ = 0;       // log2 of its entries
  uint32_t shortcut_capacity_bits_ = 0;
  uint64_t shortcut_size_ = 0;       // insert count when it was built
  bool shortcut_on_ = true;
  uint32_t multi_grid_percent_ = 100;
  uint32_t multi_mode_ = 1;
  uint32_t* scan_stack_ = nullptr;
  uint64_t scan_stack_words_ = 0;
};

}  // namespace GpuMasstree
