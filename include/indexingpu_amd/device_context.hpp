// Device-side allocator and reclaimer contexts of the device API, in the global
// namespace with the reference's names (include/device_context.hpp:18-21 declares the
// primary templates there; include/simple_slab_alloc.hpp:96-220,
// include/simple_slab_linear_alloc.hpp:125-337 and include/simple_debra_reclaim.hpp:112-318
// give their members).  One definition serves every allocator of this library (they
// share the slab_pool_view device instance) and every reclaimer (retire_view).
//
// device_allocator_context<A>: per tile.  It holds the tile's claimed bitmap word
// between cooperative calls (the next allocations need no atomics) and returns the
// unused slabs when it is destroyed.
//
// device_reclaimer_context<R>: two ways to use it.
//   * (instance, tile): per tile; retired objects go to the launch's retire list, which
//     the host frees after the launch (gpu_masstree::drain()).
//   * (instance, shmem_buffer, num_active_blocks, block): per block, as the reference's;
//     the block brackets each unit of work with begin_critical_section() and
//     end_critical_section() and calls drain_all() before it exits.  Objects its tiles
//     retire go to the block's limbo bags and are freed inside the launch once the
//     global epoch has advanced three times past the block's announced epoch (see
//     reclaim.hpp), so a long-running kernel reuses what it retires.  The bags live in
//     HBM (the reclaimer's buffer, split between the blocks of the launch); the shared
//     memory buffer of the reference's signature is not needed and may be null.
#pragma once
#include <type_traits>

#include "hip_util.hpp"
#include "layout.hpp"
#include "reclaim.hpp"
#include "slab_alloc.hpp"

namespace indexingpu_amd {

// Lanes of a tile type: Tile::numThreads (HIP cooperative-groups tiles), else 16.
template <class Tile, class = void>
struct tile_lanes : std::integral_constant<u32, 16> {};
template <class Tile>
struct tile_lanes<Tile, std::void_t<decltype(Tile::numThreads)>>
    : std::integral_constant<u32, (u32)Tile::numThreads> {};

}  // namespace indexingpu_amd

template <class Allocator>
struct device_allocator_context {
  using size_type = uint32_t;
  using pointer_type = uint32_t;
  indexingpu_amd::slab_pool_view view;
  indexingpu_amd::tile_slab_cache cache;
  uint32_t width = 16;  // lanes of the owning tile

  template <class Tile>
  __device__ device_allocator_context(const indexingpu_amd::slab_pool_view& v, const Tile&)
      : view(v), width(indexingpu_amd::tile_lanes<Tile>::value) {}
  __device__ ~device_allocator_context() {
    if (cache.bits != 0 && (__lane_id() & (width - 1)) == 0)
      indexingpu_amd::dev::atomic_and(&view.bitmap[cache.word], ~cache.bits);
  }

  // One slab, uniform over the tile (every lane calls it); 0xFFFFFFFF when the pool is
  // exhausted.
  template <class Tile>
  __device__ pointer_type allocate(const Tile&) {
    constexpr uint32_t W = indexingpu_amd::tile_lanes<Tile>::value;
    if (cache.bits == 0) {
      uint32_t w, b;
      const bool ok =
          indexingpu_amd::dev::claim_word(view, (__lane_id() & (W - 1)) == 0, w, b);
      w = __shfl(w, 0, W);
      b = __shfl(b, 0, W);
      if (!__shfl((int)ok, 0, W)) return indexingpu_amd::kInvalid;
      cache.word = w;
      cache.bits = b;
    }
    const uint32_t j = (uint32_t)__builtin_ctz(cache.bits);
    cache.bits &= cache.bits - 1;
    return cache.word * 32u + j;
  }
  template <class Tile>
  __device__ void deallocate_coop(pointer_type p, const Tile&) {
    constexpr uint32_t W = indexingpu_amd::tile_lanes<Tile>::value;
    if ((__lane_id() & (W - 1)) == 0) indexingpu_amd::dev::free_slab(view, p);
  }
  // every lane frees its own slab; the count argument of the finish call is unused
  __device__ uint32_t deallocate_perlane(pointer_type p) {
    indexingpu_amd::dev::free_slab(view, p);
    return 0;
  }
  template <class Tile>
  __device__ void deallocate_perlane_finish(uint32_t, const Tile&) {}
  __device__ void* address(pointer_type p) const {
    return static_cast<void*>(view.pool + (uint64_t)p * indexingpu_amd::kSlots);
  }

  // Linear allocator only: one past the top of the pool; element i of the linear
  // region is at get_linear() - (1 + i), as in the reference.
  __device__ void* get_linear() const {
    return static_cast<void*>(view.pool + (uint64_t)view.num_slabs * indexingpu_amd::kSlots);
  }
  // Grows the linear region to at least `size` u32 elements where the active slab
  // region leaves room; returns its size afterwards (uniform over the tile).
  template <class Tile>
  __device__ size_type reallocate_linear(size_type size, const Tile&) {
    constexpr uint32_t W = indexingpu_amd::tile_lanes<Tile>::value;
    uint32_t granted = 0;
    if ((__lane_id() & (W - 1)) == 0 && view.limits) {
      uint64_t lim = __hip_atomic_load(view.limits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (true) {
        const uint32_t act = (uint32_t)lim, lin = (uint32_t)(lim >> 32);
        if (lin >= size) {
          granted = lin;
          break;
        }
        const uint64_t room = (uint64_t)(view.num_words - act) * 1024u;
        const uint32_t want = (uint32_t)(size < room ? size : room);
        if (want <= lin) {
          granted = lin;
          break;
        }
        const uint64_t nl = ((uint64_t)want << 32) | act;
        if (__hip_atomic_compare_exchange_strong(view.limits, &lim, nl, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          granted = want;
          break;
        }
      }
    }
    return __shfl(granted, 0, W);
  }
};

template <class Reclaimer>
struct device_reclaimer_context {
  using size_type = uint32_t;
  using pointer_type = uint32_t;
  static constexpr uint32_t block_size_ = 256;
  indexingpu_amd::retire_view view;
  indexingpu_amd::tile_retire_cache cache;
  uint32_t width = 16;

  // shared memory words the block form needs (none: the bags live in HBM)
  __host__ __device__ static constexpr uint32_t required_shmem_size() { return 0; }

  // per tile: retired objects go to the launch's retire list
  template <class Tile>
  __device__ device_reclaimer_context(const indexingpu_amd::retire_view& v, const Tile&)
      : view(v), width(indexingpu_amd::tile_lanes<Tile>::value) {
    view.bag = nullptr;
  }

  // per block (every thread of the block constructs it): in-kernel epochs.  A launch
  // with more blocks than the reclaimer has announce entries, or too many to give each
  // block bags of at least 16 entries, falls back to the launch list in every block.
  template <class Block>
  __device__ device_reclaimer_context(const indexingpu_amd::retire_view& v, size_type*,
                                      size_type, const Block&)
      : view(v), width(64) {
    view.bag = nullptr;
    const uint32_t nb = gridDim.x;
    if (!v.enabled || !v.announce || !v.bags || nb > v.max_blocks) return;
    const uint32_t per = v.bag_words / nb;
    if (per < indexingpu_amd::kBagHeader + 3 * 16) return;
    view.bag_cap = (per - indexingpu_amd::kBagHeader) / 3;
    view.bag = v.bags + (uint64_t)blockIdx.x * per;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t k = 0; k < indexingpu_amd::kBagHeader; k++) view.bag[k] = 0;
      indexingpu_amd::dev::coherent_store(&view.announce[blockIdx.x], 0u);
    }
    __syncthreads();
  }

  // Retires object p (a record head, or a node slab | kRetiredNode); tile-uniform.
  template <class Tile>
  __device__ void retire(const pointer_type& p, const Tile&) {
    constexpr uint32_t W = indexingpu_amd::tile_lanes<Tile>::value;
    if (!view.enabled || (__lane_id() & (W - 1)) != 0) return;
    if (view.bag) {
      const uint32_t cur = view.bag[indexingpu_amd::kBagCurrent];
      const uint32_t k = indexingpu_amd::dev::atomic_add(&view.bag[cur], 1u);
      if (k < view.bag_cap) {
        view.bag[indexingpu_amd::kBagHeader + cur * view.bag_cap + k] = p;
        return;
      }
    }
    const uint32_t k = indexingpu_amd::dev::atomic_add(view.count, 1u);
    if (k < view.capacity) view.list[k] = p;
    else indexingpu_amd::dev::atomic_add(view.leaked, 1u);
  }

  // Every thread of the block calls it before a unit of work (block form).
  template <class Block, class Alloc>
  __device__ void begin_critical_section(const Block&, Alloc& allocator) {
    using namespace indexingpu_amd;
    __shared__ uint32_t s_epoch, s_free;
    __syncthreads();
    if (!view.bag) return;
    u32* bag = view.bag;
    const u32 cap = view.bag_cap;
    if (threadIdx.x == 0) {
      const u32 e = dev::coherent_load(view.epoch);
      u32 mask = 0;
      for (u32 k = 0; k < 3; k++)
        if (bag[k] != 0 && e - bag[3 + k] >= 3) mask |= 1u << k;
      s_epoch = e;
      s_free = mask;
    }
    __syncthreads();
    const u32 e = s_epoch, mask = s_free;
    for (u32 k = 0; k < 3; k++) {
      if (!(mask & (1u << k))) continue;
      const u32 n = bag[k] < cap ? bag[k] : cap;
      for (u32 i = threadIdx.x; i < n; i += blockDim.x)
        dev::free_retired(allocator.view, bag[kBagHeader + k * cap + i]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (u32 k = 0; k < 3; k++)
        if (mask & (1u << k)) bag[k] = 0;
      // the bag of epoch e, else an empty one (non-empty bags carry distinct epochs of
      // e - 2 .. e, so one of the three is empty when none carries e)
      u32 cur = 3;
      for (u32 k = 0; k < 3; k++)
        if (bag[k] != 0 && bag[3 + k] == e) cur = k;
      for (u32 k = 0; cur == 3 && k < 3; k++)
        if (bag[k] == 0) cur = k;
      bag[3 + cur] = e;
      bag[kBagCurrent] = cur;
      dev::coherent_store(&view.announce[blockIdx.x], (e << 1) | 1u);
      dev::order();
    }
    // advance the epoch once every block inside a critical section has announced it
    bool ok = true;
    for (u32 i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
      const u32 a = dev::coherent_load(&view.announce[i]);
      if ((a & 1u) && (a >> 1) != e) ok = false;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0 && ok) {
      u32 ex = e;
      __hip_atomic_compare_exchange_strong(view.epoch, &ex, e + 1, __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }

  // Every thread of the block calls it after a unit of work (block form).
  template <class Block>
  __device__ void end_critical_section(const Block&) {
    __syncthreads();
    if (view.bag && threadIdx.x == 0) {
      indexingpu_amd::dev::order();
      indexingpu_amd::dev::atomic_and(&view.announce[blockIdx.x], ~1u);
    }
    __syncthreads();
  }

  // Before the block exits (block form): what is left in its bags goes to the launch's
  // retire list, which the host frees after the launch.  Only the first wave of the
  // block is needed; the others return.
  template <class Block, class Tile, class Alloc>
  __device__ void drain_all(const Block&, const Tile&, Alloc&) {
    using namespace indexingpu_amd;
    if (!view.bag || threadIdx.x >= 64) return;
    u32* bag = view.bag;
    const u32 cap = view.bag_cap;
    for (u32 k = 0; k < 3; k++) {
      const u32 n = bag[k] < cap ? bag[k] : cap;
      if (n == 0) continue;
      u32 base = 0;
      if (threadIdx.x == 0) base = dev::atomic_add(view.count, n);
      base = __shfl(base, 0, 64);
      for (u32 i = threadIdx.x; i < n; i += 64) {
        if (base + i < view.capacity) view.list[base + i] = bag[kBagHeader + k * cap + i];
        else dev::atomic_add(view.leaked, 1u);
      }
    }
    if (threadIdx.x == 0) {
      bag[0] = bag[1] = bag[2] = 0;
      dev::order();
      dev::atomic_and(&view.announce[blockIdx.x], ~1u);
    }
  }
};

namespace indexingpu_amd {
using ::device_allocator_context;
using ::device_reclaimer_context;
}  // namespace indexingpu_amd
