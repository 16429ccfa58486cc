// Test library (loaded only by tests/test_gpu_reference_spelling.py): a caller written
// against the reference's own spelling of the GPU Masstree API (hyoungjook/IndexinGPU
// include/gpu_masstree.hpp, include/simple_slab_alloc.hpp,
// include/simple_slab_linear_alloc.hpp, include/simple_debra_reclaim.hpp,
// include/device_context.hpp), compiled against this repository's include/ directory
// only.  It uses the reference's include names, its global allocator / reclaimer /
// context names, GpuMasstree::gpu_masstree with tiles of 16 and 32 lanes, the template
// switches of its host calls (insert<update_if_exists, enable_suffix, use_shmem_key>,
// the erase merge-level dispatch of evaluation/gpu_masstree_adapter.hpp:215-237,
// scan<use_upper_key, concurrent>, mixed_batch<enable_suffix, ...>), and user kernels
// shaped like the reference's batch kernel (include/kernels.hpp:31-79): HIP
// cooperative-groups tiles, per-tile allocator contexts, and a per-block reclaimer
// context with begin_critical_section / end_critical_section / drain_all around each
// unit of work, calling cooperative_find / _insert / _erase / _scan.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <gpu_masstree.hpp>
#include <simple_debra_reclaim.hpp>
#include <simple_slab_alloc.hpp>
#include <simple_slab_linear_alloc.hpp>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <type_traits>

namespace cg = cooperative_groups;

namespace {

using size_type = uint32_t;
using request_type = kernels::request_type;

// ------------------------------------------------------------------ user kernels
// one find / insert / erase per tile at a time, with the per-block reclaimer form
template <typename index_type, uint32_t tile_size>
__global__ void __launch_bounds__(256) user_point_kernel(index_type index, uint32_t op,
                                                         const uint8_t* types,
                                                         const uint32_t* keys, uint32_t maxK,
                                                         const uint32_t* klens,
                                                         const uint32_t* values, uint32_t maxV,
                                                         const uint32_t* vlens, uint32_t* out,
                                                         uint32_t* out_lens, uint8_t* results,
                                                         uint32_t n) {
  using allocator_type = typename index_type::device_allocator_context_type;
  using reclaimer_type = typename index_type::device_reclaimer_context_type;
  auto block = cg::this_thread_block();
  auto tile = cg::tiled_partition<tile_size>(block);
  auto warp = cg::tiled_partition<64>(block);
  allocator_type allocator{index.allocator_, tile};
  reclaimer_type reclaimer{index.reclaimer_, nullptr, gridDim.x, block};
  const uint32_t tiles_per_block = blockDim.x / tile_size;
  const uint32_t rounds_per_block = (n + tiles_per_block - 1) / tiles_per_block;
  for (uint32_t r = blockIdx.x; r < rounds_per_block; r += gridDim.x) {
    reclaimer.begin_critical_section(block, allocator);
    const uint32_t i = r * tiles_per_block + threadIdx.x / tile_size;
    if (i < n) {
      const uint32_t* k = keys + (uint64_t)i * maxK;
      const uint32_t kl = klens[i];
      const uint32_t kind = types ? types[i] : op;
      if (kind == 2) {
        uint32_t* o = out + (uint64_t)i * maxV;
        const uint32_t len = index.template cooperative_find<true>(k, kl, o, maxV, tile, allocator);
        if (tile.thread_rank() == 0) out_lens[i] = len;
      } else if (kind == 0) {
        const uint32_t* v = values + (uint64_t)i * maxV;
        const bool ok = index.cooperative_insert(k, kl, v, vlens[i], tile, allocator, reclaimer);
        if (tile.thread_rank() == 0) results[i] = ok;
      } else {
        const bool ok = index.template cooperative_erase<true, true, true, true>(
            k, kl, tile, allocator, reclaimer);
        if (tile.thread_rank() == 0) results[i] = ok;
      }
    }
    reclaimer.end_critical_section(block);
  }
  reclaimer.drain_all(block, warp, allocator);
}

// one range query per tile at a time
template <typename index_type, uint32_t tile_size, bool use_upper_key>
__global__ void __launch_bounds__(256) user_scan_kernel(
    index_type index, const uint32_t* lower, const uint32_t* lower_lens, const uint32_t* upper,
    const uint32_t* upper_lens, uint32_t maxK, uint32_t max_count, uint32_t* counts,
    uint32_t* values, uint32_t maxV, uint32_t* vlens, uint32_t* out_keys, uint32_t* out_klens,
    uint32_t nq) {
  using allocator_type = typename index_type::device_allocator_context_type;
  auto block = cg::this_thread_block();
  auto tile = cg::tiled_partition<tile_size>(block);
  allocator_type allocator{index.allocator_, tile};
  const uint32_t tiles_per_block = blockDim.x / tile_size;
  const uint32_t ntiles = gridDim.x * tiles_per_block;
  for (uint32_t q = blockIdx.x * tiles_per_block + threadIdx.x / tile_size; q < nq; q += ntiles) {
    const uint32_t* lk = lower + (uint64_t)q * maxK;
    const uint64_t row = (uint64_t)q * max_count;
    const uint32_t c = index.template cooperative_scan<use_upper_key, true>(
        lk, lower_lens[q], tile, allocator, use_upper_key ? upper + (uint64_t)q * maxK : nullptr,
        use_upper_key ? upper_lens[q] : 1u, max_count, values + row * maxV, vlens + row, maxV,
        out_keys + row * maxK, out_klens + row, maxK);
    if (tile.thread_rank() == 0) counts[q] = c;
  }
}

// ------------------------------------------------------------------ host side
struct base {
  virtual ~base() = default;
  virtual void insert(const uint32_t*, uint32_t, const uint32_t*, const uint32_t*, uint32_t,
                      const uint32_t*, uint32_t, int, int) = 0;
  virtual void erase(const uint32_t*, uint32_t, const uint32_t*, uint32_t, uint32_t) = 0;
  virtual void find(const uint32_t*, uint32_t, const uint32_t*, uint32_t*, uint32_t, uint32_t*,
                    uint32_t, int) = 0;
  virtual void scan(const uint32_t*, const uint32_t*, uint32_t, uint32_t, uint32_t,
                    const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t*,
                    uint32_t*, uint32_t*) = 0;
  virtual void mixed(const uint8_t*, const uint32_t*, uint32_t, const uint32_t*, uint32_t*,
                     uint32_t, uint32_t*, bool*, uint32_t, int) = 0;
  virtual void user_point(uint32_t, const uint8_t*, const uint32_t*, uint32_t, const uint32_t*,
                          const uint32_t*, uint32_t, const uint32_t*, uint32_t*, uint32_t*,
                          uint8_t*, uint32_t, uint32_t, int) = 0;
  virtual uint32_t pending_retired() = 0;
  virtual void drain() = 0;
  virtual void user_scan(const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*,
                         uint32_t, uint32_t, uint32_t*, uint32_t*, uint32_t, uint32_t*, uint32_t*,
                         uint32_t*, uint32_t) = 0;
  virtual uint64_t slabs_used() = 0;
  virtual void stats() = 0;
  virtual uint32_t error() = 0;
};

template <typename allocator_type, uint32_t tile_size>
struct holder final : base {
  using reclaimer_type = simple_debra_reclaimer<>;
  using index_type = GpuMasstree::gpu_masstree<allocator_type, reclaimer_type, tile_size>;
  std::unique_ptr<allocator_type> alloc;
  std::unique_ptr<reclaimer_type> recl;
  std::unique_ptr<index_type> index;
  explicit holder(float pool_ratio) {
    // the linear allocator starts with a small active slab region, so the tests grow it
    if constexpr (std::is_same<allocator_type, simple_slab_linear_allocator<128>>::value)
      alloc.reset(new allocator_type(pool_ratio, 0.02f));
    else
      alloc.reset(new allocator_type(pool_ratio));
    recl.reset(new reclaimer_type());
    index.reset(new index_type(*alloc, *recl));
  }
  ~holder() override {
    (void)hipDeviceSynchronize();
    index.reset();
    recl.reset();
    alloc.reset();
  }

  template <bool enable_suffix, typename... A>
  void do_insert(int upsert, A... a) {
    if (upsert) index->template insert<true, enable_suffix, false>(a...);
    else index->template insert<false, enable_suffix, false>(a...);
  }
  void insert(const uint32_t* k, uint32_t mk, const uint32_t* kl, const uint32_t* v, uint32_t mv,
              const uint32_t* vl, uint32_t n, int upsert, int enable_suffix) override {
    if (enable_suffix) do_insert<true>(upsert, k, mk, kl, v, mv, vl, n, hipStream_t(0));
    else do_insert<false>(upsert, k, mk, kl, v, mv, vl, n, hipStream_t(0));
  }
  // evaluation/gpu_masstree_adapter.hpp:215-219
  template <uint32_t merge_level, typename... A>
  void do_erase(A... a) {
    index->template erase<merge_level >= 4, merge_level >= 3, merge_level >= 2, merge_level >= 1,
                          false>(a...);
  }
  void erase(const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t n,
             uint32_t merge_level) override {
    switch (merge_level) {
      case 0: do_erase<0>(k, mk, kl, n, hipStream_t(0)); break;
      case 1: do_erase<1>(k, mk, kl, n, hipStream_t(0)); break;
      case 2: do_erase<2>(k, mk, kl, n, hipStream_t(0)); break;
      case 3: do_erase<3>(k, mk, kl, n, hipStream_t(0)); break;
      default: do_erase<4>(k, mk, kl, n, hipStream_t(0)); break;
    }
  }
  void find(const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t* v, uint32_t mv,
            uint32_t* vl, uint32_t n, int concurrent) override {
    if (concurrent) index->template find<true, false>(k, mk, kl, v, mv, vl, n);
    else index->template find<false, false>(k, mk, kl, v, mv, vl, n);
  }
  void scan(const uint32_t* lo, const uint32_t* lol, uint32_t mk, uint32_t maxc, uint32_t nq,
            const uint32_t* up, const uint32_t* upl, uint32_t* counts, uint32_t* v, uint32_t mv,
            uint32_t* vl, uint32_t* ok, uint32_t* okl) override {
    if (up) index->template scan<true, false, false>(lo, lol, mk, maxc, nq, up, upl, counts, v, mv, vl, ok, okl);
    else index->template scan<false, false, false>(lo, lol, mk, maxc, nq, nullptr, nullptr, counts, v, mv, vl, ok, okl);
  }
  void mixed(const uint8_t* t, const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t* v,
             uint32_t mv, uint32_t* vl, bool* r, uint32_t n, int merge_level) override {
    auto types = reinterpret_cast<const request_type*>(t);
    // evaluation/gpu_masstree_adapter.hpp:233-237
    if (merge_level >= 2) index->template mixed_batch<true, true, true, true, false>(types, k, mk, kl, v, mv, vl, r, n);
    else index->template mixed_batch<true, false, false, false, false>(types, k, mk, kl, v, mv, vl, r, n);
  }
  void user_point(uint32_t op, const uint8_t* types, const uint32_t* k, uint32_t mk,
                  const uint32_t* kl, const uint32_t* v, uint32_t mv, const uint32_t* vl,
                  uint32_t* out, uint32_t* outl, uint8_t* res, uint32_t n, uint32_t blocks,
                  int drain) override {
    hipLaunchKernelGGL((user_point_kernel<index_type, tile_size>), dim3(blocks), dim3(256), 0, 0,
                       *index, op, types, k, mk, kl, v, mv, vl, out, outl, res, n);
    if (drain) index->drain();  // what drain_all handed to the launch list
  }
  uint32_t pending_retired() override { return recl->pending(); }
  void drain() override { index->drain(); }
  void user_scan(const uint32_t* lo, const uint32_t* lol, const uint32_t* up, const uint32_t* upl,
                 uint32_t mk, uint32_t maxc, uint32_t* counts, uint32_t* v, uint32_t mv,
                 uint32_t* vl, uint32_t* ok, uint32_t* okl, uint32_t nq) override {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(1024, (nq + 15) / 16));
    if (up)
      hipLaunchKernelGGL((user_scan_kernel<index_type, tile_size, true>), dim3(blocks), dim3(256),
                         0, 0, *index, lo, lol, up, upl, mk, maxc, counts, v, mv, vl, ok, okl, nq);
    else
      hipLaunchKernelGGL((user_scan_kernel<index_type, tile_size, false>), dim3(blocks), dim3(256),
                         0, 0, *index, lo, lol, up, upl, mk, maxc, counts, v, mv, vl, ok, okl, nq);
  }
  uint64_t slabs_used() override { return index->num_nodes_used(); }
  void stats() override { alloc->print_stats(); }
  uint32_t error() override { return index->last_error(); }
};

}  // namespace

extern "C" {

// linear: 0 = simple_slab_allocator, 1 = simple_slab_linear_allocator
void* rs_create(uint32_t tile, uint32_t linear, float pool_ratio) {
  try {
    if (linear) {
      if (tile == 32) return new holder<simple_slab_linear_allocator<128>, 32>(pool_ratio);
      return new holder<simple_slab_linear_allocator<128>, 16>(pool_ratio);
    }
    if (tile == 32) return new holder<simple_slab_allocator<128>, 32>(pool_ratio);
    return new holder<simple_slab_allocator<128>, 16>(pool_ratio);
  } catch (...) {
    return nullptr;
  }
}
void rs_destroy(void* h) { delete static_cast<base*>(h); }
uint32_t rs_error(void* h) { return static_cast<base*>(h)->error(); }
uint64_t rs_slabs_used(void* h) { return static_cast<base*>(h)->slabs_used(); }
uint32_t rs_launch_error(void) { return (uint32_t)hipGetLastError(); }
void rs_stats(void* h) {
  static_cast<base*>(h)->stats();
  std::fflush(stdout);
}
void rs_insert(void* h, const uint32_t* k, uint32_t mk, const uint32_t* kl, const uint32_t* v,
               uint32_t mv, const uint32_t* vl, uint32_t n, int upsert, int enable_suffix) {
  static_cast<base*>(h)->insert(k, mk, kl, v, mv, vl, n, upsert, enable_suffix);
}
void rs_erase(void* h, const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t n,
              uint32_t merge_level) {
  static_cast<base*>(h)->erase(k, mk, kl, n, merge_level);
}
void rs_find(void* h, const uint32_t* k, uint32_t mk, const uint32_t* kl, uint32_t* v,
             uint32_t mv, uint32_t* vl, uint32_t n, int concurrent) {
  static_cast<base*>(h)->find(k, mk, kl, v, mv, vl, n, concurrent);
}
void rs_scan(void* h, const uint32_t* lo, const uint32_t* lol, uint32_t mk, uint32_t maxc,
             uint32_t nq, const uint32_t* up, const uint32_t* upl, uint32_t* counts, uint32_t* v,
             uint32_t mv, uint32_t* vl, uint32_t* ok, uint32_t* okl) {
  static_cast<base*>(h)->scan(lo, lol, mk, maxc, nq, up, upl, counts, v, mv, vl, ok, okl);
}
void rs_mixed(void* h, const uint8_t* t, const uint32_t* k, uint32_t mk, const uint32_t* kl,
              uint32_t* v, uint32_t mv, uint32_t* vl, bool* r, uint32_t n, int merge_level) {
  static_cast<base*>(h)->mixed(t, k, mk, kl, v, mv, vl, r, n, merge_level);
}
// op: 0 insert, 1 erase, 2 find (user kernel, per-block reclaimer context)
// types (optional): per-request kind, overriding op.  drain = 0 leaves what drain_all
// handed to the launch's retire list in place (rs_pending_retired counts it).
void rs_user_point(void* h, uint32_t op, const uint8_t* types, const uint32_t* k, uint32_t mk,
                   const uint32_t* kl, const uint32_t* v, uint32_t mv, const uint32_t* vl,
                   uint32_t* out, uint32_t* outl, uint8_t* res, uint32_t n, uint32_t blocks,
                   int drain) {
  static_cast<base*>(h)->user_point(op, types, k, mk, kl, v, mv, vl, out, outl, res, n, blocks,
                                    drain);
}
uint32_t rs_pending_retired(void* h) { return static_cast<base*>(h)->pending_retired(); }
void rs_drain(void* h) { static_cast<base*>(h)->drain(); }
void rs_user_scan(void* h, const uint32_t* lo, const uint32_t* lol, const uint32_t* up,
                  const uint32_t* upl, uint32_t mk, uint32_t maxc, uint32_t* counts, uint32_t* v,
                  uint32_t mv, uint32_t* vl, uint32_t* ok, uint32_t* okl, uint32_t nq) {
  static_cast<base*>(h)->user_scan(lo, lol, up, upl, mk, maxc, counts, v, mv, vl, ok, okl, nq);
}

}  // extern "C"
