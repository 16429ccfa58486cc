// Safe reclamation for the MI355X Masstree.
//
// Host objects keep the reference's names so the template signature is a drop-in:
// simple_debra_reclaimer<buffer_size_per_block> and simple_dummy_reclaimer
// (include/simple_debra_reclaim.hpp:32-110, include/simple_dummy_reclaim.hpp:26-64).
//
// Mechanism (this project's design).  A retired object (a record, or a node that an
// erase unlinked) may still be read by an optimistic reader of the same launch, so it
// is not freed in place.
//   * Batch kernels: the retiring tile appends it to a device retire list, reserving 16
//     entries at a time with one atomic.  After the launch, the host enqueues a drain
//     kernel on the same stream; stream order guarantees every reader of that launch
//     has finished, so the drain frees each retired object and resets the list.  This
//     is epoch-based reclamation with the launch as the epoch: no per-iteration
//     announce exchange, no block barriers inside the batch kernels.  A full list leaks
//     the object (counted in `leaked`) instead of blocking.
//   * User kernels of the device API that bracket their work with the reclaimer
//     context's begin_critical_section / end_critical_section (device_context.hpp):
//     DEBRA-style epochs inside the launch, so a long-running kernel reuses the slabs it
//     retires.  Each block announces the global epoch it observed; a block whose
//     announced epoch is current everywhere advances it; objects a block retired while
//     announcing epoch e are freed by that block once it observes epoch e + 3 (three
//     limbo bags per block).  drain_all hands what is left to the launch list.
// Launches that mutate one index must be ordered on one stream, as the reference
// requires for its announce array (include/simple_debra_reclaim.hpp:189,226).
#pragma once
#include <cstdint>

#include "hip_util.hpp"
#include "layout.hpp"

namespace indexingpu_amd {

struct retire_view {
  u32* list;      // capacity entries (record head indices, kInvalid = padding)
  u32* count;     // entries reserved so far this epoch
  u32* leaked;    // records dropped because the list was full
  u32 capacity;
  u32 enabled;    // 0: retire is a no-op (simple_dummy_reclaimer)
  // In-kernel epoch reclamation for user kernels of the device API (the context's
  // begin_critical_section / end_critical_section / drain_all; device_context.hpp).
  u32* epoch;     // global epoch
  u32* announce;  // per block of the launch: epoch << 1 | 1 while in a critical section
  u32* bags;      // limbo bags, split evenly between the blocks of a launch
  u32 bag_words;  // words of `bags`
  u32 max_blocks; // entries of `announce`
  // Set by a block's reclaimer context for its own tiles (null in the batch kernels):
  // this block's region of `bags` (header of kBagHeader words, then 3 bags of bag_cap
  // entries); retired objects go to the current bag instead of the launch list.
  u32* bag;
  u32 bag_cap;
};
constexpr u32 kBagHeader = 8;  // counts[3], epoch tags[3], -, current bag
constexpr u32 kBagCurrent = 7;

template <uint32_t buffer_size_per_block = 131072>
class simple_debra_reclaimer {
 public:
  using size_type = uint32_t;
  using device_instance_type = retire_view;
  static constexpr uint32_t block_size_ = 256;
  static constexpr uint32_t kMaxBlocks = 1u << 16;
  static constexpr uint32_t kBagBlocks = 256;

  simple_debra_reclaimer() {
    int dev = 0, cus = 0;
    IGPU_HIP_TRY(hipGetDevice(&dev));
    IGPU_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    uint64_t cap = (uint64_t)(buffer_size_per_block / sizeof(u32)) * (uint64_t)cus * 8ull;
    if (cap > (1ull << 26)) cap = 1ull << 26;
    if (cap < 4096) cap = 4096;
    init((u32)cap);
  }
  explicit simple_debra_reclaimer(uint32_t capacity_entries) { init(capacity_entries); }
  ~simple_debra_reclaimer() {
    if (view_.list) (void)hipFree(view_.list);
    if (view_.count) (void)hipFree(view_.count);
    if (view_.announce) (void)hipFree(view_.announce);
    if (view_.bags) (void)hipFree(view_.bags);
  }
  simple_debra_reclaimer(const simple_debra_reclaimer&) = delete;
  simple_debra_reclaimer& operator=(const simple_debra_reclaimer&) = delete;

  device_instance_type get_device_instance() const { return view_; }

  uint32_t leaked() const {
    u32 h = 0;
    IGPU_HIP_TRY(hipMemcpy(&h, view_.leaked, 4, hipMemcpyDeviceToHost));
    return h;
  }
  // entries in the launch retire list not yet drained (synchronizes)
  uint32_t pending() const {
    u32 h = 0;
    IGPU_HIP_TRY(hipDeviceSynchronize());
    IGPU_HIP_TRY(hipMemcpy(&h, view_.count, 4, hipMemcpyDeviceToHost));
    return h;
  }

 private:
  void init(u32 cap) {
    view_.capacity = cap;
    view_.enabled = 1;
    IGPU_HIP_TRY(hipMalloc(&view_.list, sizeof(u32) * (size_t)cap));
    IGPU_HIP_TRY(hipMalloc(&view_.count, 256));
    IGPU_HIP_TRY(hipMemset(view_.count, 0, 256));
    view_.leaked = view_.count + 1;
    view_.epoch = view_.count + 32;  // its own 128-byte line
    // device-API epochs: announce entries for up to kMaxBlocks blocks per launch and
    // buffer_size_per_block bytes of limbo bags for each of up to kBagBlocks blocks
    view_.max_blocks = kMaxBlocks;
    IGPU_HIP_TRY(hipMalloc(&view_.announce, 4ull * kMaxBlocks));
    IGPU_HIP_TRY(hipMemset(view_.announce, 0, 4ull * kMaxBlocks));
    view_.bag_words = (u32)((uint64_t)buffer_size_per_block / 4 * kBagBlocks);
    IGPU_HIP_TRY(hipMalloc(&view_.bags, 4ull * view_.bag_words));
    IGPU_HIP_TRY(hipMemset(view_.bags, 0, 4ull * view_.bag_words));
    IGPU_HIP_TRY(hipDeviceSynchronize());
  }
  retire_view view_{};
};

// No reclamation: retired records are simply leaked (the reference's dummy reclaimer).
class simple_dummy_reclaimer {
 public:
  using size_type = uint32_t;
  using device_instance_type = retire_view;
  static constexpr uint32_t block_size_ = 256;
  simple_dummy_reclaimer() = default;
  device_instance_type get_device_instance() const { return retire_view{}; }
  uint32_t leaked() const { return 0; }
};

// Per-tile reservation in the retire list (uniform over the tile).
struct tile_retire_cache {
  u32 base = kInvalid;
  u32 used = 0;
};
constexpr u32 kRetireChunk = 16;

}  // namespace indexingpu_amd

// The reference declares its reclaimers in the global namespace
// (include/simple_debra_reclaim.hpp:32-33, include/simple_dummy_reclaim.hpp:26-27).
using indexingpu_amd::simple_debra_reclaimer;
using indexingpu_amd::simple_dummy_reclaimer;
