// Slab allocator for the MI355X Masstree: one HBM pool of 128-byte slabs (one gfx950
// L2 line each) with a 1-bit-per-slab occupancy bitmap.
//
// Host object: same name, template parameter and constructor as the reference's
// simple_slab_allocator<slab_size>(float pool_ratio = 0.9f), where the pool is
// pool_ratio x total device memory (include/simple_slab_alloc.hpp:28-94), plus a
// constructor taking an explicit byte size.  print_stats() prints the reference's line
// "simple_slab_allocator(128B slabs): A/T slabs allocated (P%)" (:82-89).
//
// Device side (this project's design, not the reference's per-tile hashed probing):
// a tile claims a whole 32-slab bitmap word with one atomic fetch_or(0xFFFFFFFF) at a
// round-robin cursor and keeps the word's free bits in registers; the next 31
// allocations of that tile need no atomics at all.  Unused bits go back with one
// fetch_and when the tile finishes.  Frees clear single bits.
#pragma once
#include <cstdint>
#include <algorithm>
#include <iostream>
#include <utility>
#include <vector>

#include "hip_util.hpp"
#include "layout.hpp"

namespace indexingpu_amd {

// POD view of the pool passed to kernels.
struct slab_pool_view {
  u64* pool;    // num_slabs * 16 slots
  u32* bitmap;  // num_words words, bit b of word w = slab 32w + b allocated
  u32* cursor;  // round-robin claim cursor (in bitmap words)
  u32 num_slabs;
  u32 num_words;
  // Linear allocator only (null otherwise): low 32 bits = bitmap words of the active
  // slab region [0, active), high 32 bits = u32 elements of the linear region at the top
  // of the pool.  Claims stay inside the active region, which grows when claims find it
  // nearly full; the two regions never overlap.
  u64* limits;
  u32 grow_words;  // growth step of the active region (bitmap words)
  u32 no_reuse;    // 1: frees are ignored (bump allocator)

  std::size_t num_allocated_slabs() const {
    std::vector<u32> h(num_words);
    IGPU_HIP_TRY(hipMemcpy(h.data(), bitmap, sizeof(u32) * num_words, hipMemcpyDeviceToHost));
    std::size_t c = 0;
    for (u32 w : h) c += (std::size_t)__builtin_popcount(w);
    return c;
  }
};

struct pool_bytes {
  std::size_t bytes;
};

template <uint32_t slab_size = 128>
class simple_slab_allocator {
 public:
  static_assert(slab_size == kSlabBytes, "the node layout is built on 128-byte slabs");
  using size_type = uint32_t;
  using pointer_type = uint32_t;
  static constexpr uint32_t slab_size_ = slab_size;
  using device_instance_type = slab_pool_view;

  explicit simple_slab_allocator(float pool_ratio = 0.9f) {
    size_t free_b = 0, total_b = 0;
    IGPU_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    init(static_cast<std::size_t>(static_cast<double>(total_b) * pool_ratio));
  }
  explicit simple_slab_allocator(pool_bytes b) { init(b.bytes); }
  ~simple_slab_allocator() {
    if (base_) (void)hipFree(base_);
  }
  simple_slab_allocator(const simple_slab_allocator&) = delete;
  simple_slab_allocator& operator=(const simple_slab_allocator&) = delete;

  device_instance_type get_device_instance() const { return view_; }

  std::size_t num_allocated_slabs() const { return view_.num_allocated_slabs(); }
  uint64_t total_slabs() const { return view_.num_slabs; }

  void print_stats() const {
    auto used = num_allocated_slabs();
    std::cout << "simple_slab_allocator(" << slab_size_ << "B slabs): " << used << "/"
              << view_.num_slabs << " slabs allocated ("
              << (float)used / (float)view_.num_slabs * 100.0f << "%)" << std::endl;
  }

 private:
  void init(std::size_t bytes) {
    uint64_t slabs = bytes / kSlabBytes;
    slabs = slabs / 32 * 32;  // whole bitmap words
    // slab indices stay below 2^31: bit 31 of an interior child index is the version tag
    if (slabs > kMaxPoolSlabs) slabs = kMaxPoolSlabs;
    if (slabs < 64) slabs = 64;
    view_.num_slabs = (u32)slabs;
    view_.num_words = (u32)(slabs / 32);
    std::size_t bitmap_bytes = (std::size_t)view_.num_words * 4;
    std::size_t head = (bitmap_bytes + 256 + 255) / 256 * 256;  // bitmap + cursor, aligned
    std::size_t total = head + (std::size_t)slabs * kSlabBytes;
    IGPU_HIP_TRY(hipMalloc(&base_, total));
    IGPU_HIP_TRY(hipMemset(base_, 0, head));
    view_.bitmap = reinterpret_cast<u32*>(base_);
    view_.cursor = reinterpret_cast<u32*>(reinterpret_cast<char*>(base_) + bitmap_bytes);
    view_.pool = reinterpret_cast<u64*>(reinterpret_cast<char*>(base_) + head);
    IGPU_HIP_TRY(hipDeviceSynchronize());
  }

  void* base_ = nullptr;
  slab_pool_view view_{};
};

// Linear-region allocator: same name, template parameter and constructor as the
// reference's simple_slab_linear_allocator<slab_size>(float pool_ratio = 0.9f,
// float initial_slab_ratio = 0.7f) (include/simple_slab_linear_alloc.hpp:31-123), and the
// same print_stats() line.  Slab claims are confined to an active prefix of the bitmap
// (initially initial_slab_ratio of the pool, in whole growth steps) that grows by
// 1/64 of the pool whenever claims sweep past many full words, so a tree's nodes stay
// packed in a small, TLB-friendly region; the rest of the pool above the active region
// can be given to a linear array of u32 elements that grows down from the top of the
// pool (device_allocator_context::reallocate_linear / get_linear).
template <uint32_t slab_size = 128>
class simple_slab_linear_allocator {
 public:
  static_assert(slab_size == kSlabBytes, "the node layout is built on 128-byte slabs");
  using size_type = uint32_t;
  using pointer_type = uint32_t;
  static constexpr uint32_t slab_size_ = slab_size;
  using device_instance_type = slab_pool_view;

  explicit simple_slab_linear_allocator(float pool_ratio = 0.9f, float initial_slab_ratio = 0.7f) {
    size_t free_b = 0, total_b = 0;
    IGPU_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    init(static_cast<std::size_t>(static_cast<double>(total_b) * pool_ratio), initial_slab_ratio);
  }
  simple_slab_linear_allocator(pool_bytes b, float initial_slab_ratio = 0.7f) {
    init(b.bytes, initial_slab_ratio);
  }
  ~simple_slab_linear_allocator() {
    if (base_) (void)hipFree(base_);
  }
  simple_slab_linear_allocator(const simple_slab_linear_allocator&) = delete;
  simple_slab_linear_allocator& operator=(const simple_slab_linear_allocator&) = delete;

  device_instance_type get_device_instance() const { return view_; }
  std::size_t num_allocated_slabs() const { return view_.num_allocated_slabs(); }
  uint64_t total_slabs() const { return view_.num_slabs; }
  // (active bitmap words, linear elements)
  std::pair<uint32_t, uint32_t> limits() const {
    u64 h = 0;
    IGPU_HIP_TRY(hipMemcpy(&h, view_.limits, 8, hipMemcpyDeviceToHost));
    return {(uint32_t)h, (uint32_t)(h >> 32)};
  }

  void print_stats() const {
    const auto used = num_allocated_slabs();
    const auto lim = limits();
    // blocks = bitmap words of 32 slabs
    std::cout << "simple_slab_linear_allocator(" << slab_size_ << "B slabs, " << lim.first << "/"
              << view_.num_words << " blocks): " << used << "/" << view_.num_slabs
              << " slabs + " << lim.second << " linear allocated ("
              << (float)((double)used * slab_size_ + (double)lim.second * 4.0) /
                     ((double)view_.num_slabs * slab_size_) * 100.0f
              << "%)" << std::endl;
  }

 private:
  void init(std::size_t bytes, float initial_slab_ratio) {
    uint64_t slabs = bytes / kSlabBytes;
    slabs = slabs / 32 * 32;
    if (slabs > kMaxPoolSlabs) slabs = kMaxPoolSlabs;
    if (slabs < 64 * 32) slabs = 64 * 32;
    view_.num_slabs = (u32)slabs;
    view_.num_words = (u32)(slabs / 32);
    view_.grow_words = std::max<u32>(1, view_.num_words / 64);
    std::size_t bitmap_bytes = (std::size_t)view_.num_words * 4;
    std::size_t head = (bitmap_bytes + 512 + 255) / 256 * 256;  // bitmap, cursor, limits
    std::size_t total = head + (std::size_t)slabs * kSlabBytes;
    IGPU_HIP_TRY(hipMalloc(&base_, total));
    IGPU_HIP_TRY(hipMemset(base_, 0, head));
    view_.bitmap = reinterpret_cast<u32*>(base_);
    view_.cursor = reinterpret_cast<u32*>(reinterpret_cast<char*>(base_) + bitmap_bytes);
    view_.limits = reinterpret_cast<u64*>(reinterpret_cast<char*>(base_) + bitmap_bytes + 256);
    view_.pool = reinterpret_cast<u64*>(reinterpret_cast<char*>(base_) + head);
    double r = initial_slab_ratio;
    if (!(r > 0.0)) r = 0.0;
    if (r > 1.0) r = 1.0;
    uint64_t act = (uint64_t)(r * view_.num_words);
    act = (act + view_.grow_words - 1) / view_.grow_words * view_.grow_words;
    if (act < view_.grow_words) act = view_.grow_words;
    if (act > view_.num_words) act = view_.num_words;
    const u64 lim = act;
    IGPU_HIP_TRY(hipMemcpy(view_.limits, &lim, 8, hipMemcpyHostToDevice));
    IGPU_HIP_TRY(hipDeviceSynchronize());
  }

  void* base_ = nullptr;
  slab_pool_view view_{};
};

// Allocator that never reuses a slab (frees are ignored): the reference's
// simple_bump_allocator<slab_size> name and constructor, over the same pool and device
// view as simple_slab_allocator.
template <uint32_t slab_size = 128>
class simple_bump_allocator : public simple_slab_allocator<slab_size> {
 public:
  explicit simple_bump_allocator(float pool_ratio = 0.9f)
      : simple_slab_allocator<slab_size>(pool_ratio) {}
  explicit simple_bump_allocator(pool_bytes b) : simple_slab_allocator<slab_size>(b) {}
  slab_pool_view get_device_instance() const {
    slab_pool_view v = simple_slab_allocator<slab_size>::get_device_instance();
    v.no_reuse = 1;
    return v;
  }
};

// Per-tile cache of one claimed bitmap word (uniform over the tile).
struct tile_slab_cache {
  u32 word = kInvalid;
  u32 bits = 0;
};

namespace dev {

// Linear allocator: grows the active slab region by one step unless the linear region
// above it leaves no room.  Returns true when the region is larger than `seen` words
// (grown here or by another tile).
__device__ __forceinline__ bool grow_active(const slab_pool_view& v, u32 seen) {
  u64 lim = __hip_atomic_load(v.limits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const u32 act = (u32)lim, lin = (u32)(lim >> 32);
    if (act > seen) return true;
    const u32 lin_words = (u32)(((u64)lin + 1023) / 1024);  // 1024 u32 per bitmap word
    const u32 top = v.num_words - (lin_words < v.num_words ? lin_words : v.num_words);
    if (act >= top) return false;
    const u32 na = act + v.grow_words < top ? act + v.grow_words : top;
    const u64 want = ((u64)lin << 32) | na;
    if (__hip_atomic_compare_exchange_strong(v.limits, &lim, want, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return true;
  }
}

// claims after which a linear allocator's sweep that found only full words grows the
// active region
constexpr u32 kGrowAfterFull = 32;

// Claims a bitmap word with free slabs.  Executed by every lane of the tile; only
// lane `leader` touches memory; the result is returned uniformly through the caller's
// broadcast.  Returns false when the pool (for a linear allocator: the active region,
// grown as far as the linear region allows) is exhausted.
__device__ __forceinline__ bool claim_word(const slab_pool_view& v, bool leader, u32& word,
                                           u32& bits) {
  word = kInvalid;
  bits = 0;
  if (!leader) return true;
  if (!v.limits) {
    const u32 tries = v.num_words + 1;
    for (u32 t = 0; t < tries; t++) {
      u32 w = atomic_add(v.cursor, 1u) % v.num_words;
      u32 old = atomic_or(&v.bitmap[w], 0xFFFFFFFFu);
      u32 fr = ~old;
      if (fr != 0) {
        word = w;
        bits = fr;
        return true;
      }
    }
    return false;
  }
  u32 misses = 0;
  while (true) {
    const u32 nw = (u32)__hip_atomic_load(v.limits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool grew = false;
    for (u32 t = 0; t <= nw; t++) {
      const u32 w = atomic_add(v.cursor, 1u) % nw;
      const u32 old = atomic_or(&v.bitmap[w], 0xFFFFFFFFu);
      if (old != 0xFFFFFFFFu) {
        word = w;
        bits = ~old;
        return true;
      }
      if (++misses % kGrowAfterFull == 0 && grow_active(v, nw)) {
        grew = true;
        break;
      }
    }
    // every active word was full: grow, or report the pool exhausted
    if (!grew && !grow_active(v, nw)) return false;
  }
}

__device__ __forceinline__ void free_slab(const slab_pool_view& v, u32 idx) {
  if (v.no_reuse) return;
  atomic_and(&v.bitmap[idx >> 5], ~(1u << (idx & 31)));
}

// Frees a retired object: a record chain, or (kRetiredNode set) a single node slab.
__device__ __forceinline__ void free_retired(const slab_pool_view& v, u32 r) {
  if (r & kRetiredNode) {
    const u32 idx = r & ~kRetiredNode;
    if (idx < v.num_slabs) free_slab(v, idx);
    return;
  }
  if (r >= v.num_slabs) return;
  const u32* w = reinterpret_cast<const u32*>(v.pool + (u64)r * kSlots);
  const u32 lw = w[0];
  const u32 ns = rec_num_slabs(1 + rec_tail_len(lw) + rec_value_len(lw));
  u32 cur = r;
  for (u32 s = 0; s < ns && cur < v.num_slabs; s++) {
    const u32 nxt = reinterpret_cast<const u32*>(v.pool + (u64)cur * kSlots)[kRecordNextWord];
    free_slab(v, cur);
    cur = nxt;
  }
}

}  // namespace dev
}  // namespace indexingpu_amd

// The reference declares its allocators in the global namespace
// (include/simple_slab_alloc.hpp:28-29, include/simple_slab_linear_alloc.hpp:31-32,
// include/simple_bump_alloc.hpp); callers spell them that way.
using indexingpu_amd::simple_bump_allocator;
using indexingpu_amd::simple_slab_allocator;
using indexingpu_amd::simple_slab_linear_allocator;
