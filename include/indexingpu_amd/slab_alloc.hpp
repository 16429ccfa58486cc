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
#include <iostream>
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

// Per-tile cache of one claimed bitmap word (uniform over the tile).
struct tile_slab_cache {
  u32 word = kInvalid;
  u32 bits = 0;
};

namespace dev {

// Claims a bitmap word with free slabs.  Executed by every lane of the tile; only
// lane `leader` touches memory; the result is returned uniformly through the caller's
// broadcast.  Returns false when the pool is exhausted.
__device__ __forceinline__ bool claim_word(const slab_pool_view& v, bool leader, u32& word,
                                           u32& bits) {
  word = kInvalid;
  bits = 0;
  if (!leader) return true;
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

__device__ __forceinline__ void free_slab(const slab_pool_view& v, u32 idx) {
  atomic_and(&v.bitmap[idx >> 5], ~(1u << (idx & 31)));
}

}  // namespace dev
}  // namespace indexingpu_amd
