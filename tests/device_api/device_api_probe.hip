// Test library for the device-side API (test infrastructure: loaded only by
// tests/test_gpu_device_api.py).  An index is built with the C++ host API, and a
// kernel applies each batch through gpu_masstree::cooperative_insert / _update /
// _find / _erase, one request per 16-lane tile at a time, the way the reference's
// device interface is used from user kernels (include/gpu_masstree.hpp:194-971 in
// hyoungjook/IndexinGPU): the tree is passed to the kernel by value (a shallow copy),
// and each tile builds its allocator and reclaimer contexts.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>

#include "indexingpu_amd/gpu_masstree.hpp"

namespace {

using alloc_t = indexingpu_amd::simple_slab_allocator<128>;
using reclaim_t = indexingpu_amd::simple_debra_reclaimer<>;
using tree_t = GpuMasstree::gpu_masstree<alloc_t, reclaim_t, 16>;

struct probe {
  alloc_t* alloc;
  reclaim_t* reclaim;
  tree_t* tree;
};

struct tile16 {};  // the device API takes the tile's type only

__global__ void __launch_bounds__(256) device_api_kernel(
    tree_t tree, indexingpu_amd::slab_pool_view av, indexingpu_amd::retire_view rv, uint32_t op,
    const uint32_t* keys, uint32_t maxK, const uint32_t* klens, const uint32_t* values,
    uint32_t maxV, const uint32_t* vlens, uint32_t* out_values, uint32_t* out_vlens,
    uint8_t* results, uint32_t n) {
  const tile16 tile;
  GpuMasstree::device_allocator_context<alloc_t> actx(av, tile);
  GpuMasstree::device_reclaimer_context<reclaim_t> rctx(rv, tile);
  const uint32_t lane = threadIdx.x & 15u;
  const uint64_t per_block = blockDim.x / 16;
  const uint64_t ntiles = (uint64_t)gridDim.x * per_block;
  for (uint64_t i = blockIdx.x * per_block + threadIdx.x / 16; i < n; i += ntiles) {
    const uint32_t* k = keys + i * maxK;
    const uint32_t kl = klens ? klens[i] : maxK;
    if (op == 2) {
      uint32_t* o = out_values + i * maxV;
      const uint32_t len = tree.template cooperative_find<true>(k, kl, o, maxV, tile, actx);
      if (lane == 0) out_vlens[i] = len;
    } else {
      bool r = false;
      const uint32_t* v = values ? values + i * maxV : nullptr;
      const uint32_t vl = vlens ? vlens[i] : maxV;
      if (op == 0) r = tree.cooperative_insert(k, kl, v, vl, tile, actx, rctx);
      else if (op == 4) r = tree.template cooperative_insert<true>(k, kl, v, vl, tile, actx, rctx);
      else if (op == 3) r = tree.cooperative_update(k, kl, v, vl, tile, actx, rctx);
      else r = tree.template cooperative_erase<true, false, false, false>(k, kl, tile, actx, rctx);
      if (lane == 0) results[i] = r ? 1 : 0;
    }
  }
}

}  // namespace

extern "C" {

// Creates an index with a pool of pool_bytes; null on failure.
void* dapi_create(uint64_t pool_bytes) {
  try {
    probe* p = new probe{};
    p->alloc = new alloc_t(indexingpu_amd::pool_bytes{pool_bytes});
    p->reclaim = new reclaim_t();
    p->tree = new tree_t(*p->alloc, *p->reclaim);
    return p;
  } catch (...) {
    return nullptr;
  }
}

void dapi_destroy(void* h) {
  probe* p = static_cast<probe*>(h);
  if (!p) return;
  (void)hipDeviceSynchronize();
  delete p->tree;
  delete p->reclaim;
  delete p->alloc;
  delete p;
}

// Applies n requests of kind op (0 insert, 1 erase, 2 find, 3 update, 4 upsert) through
// the device API, waits, returns retired slabs to the allocator, and returns the
// device error code (0 = none).  Device pointers throughout; null key_lengths /
// value_lengths mean full length.
uint32_t dapi_apply(void* h, uint32_t op, const uint32_t* keys, uint32_t maxK,
                    const uint32_t* klens, const uint32_t* values, uint32_t maxV,
                    const uint32_t* vlens, uint32_t* out_values, uint32_t* out_vlens,
                    uint8_t* results, uint32_t n) {
  probe* p = static_cast<probe*>(h);
  if (n) {
    const uint32_t blocks = (n + 15) / 16 < 1024 ? (n + 15) / 16 : 1024;
    hipLaunchKernelGGL(device_api_kernel, dim3(blocks), dim3(256), 0, 0, *p->tree,
                       p->alloc->get_device_instance(), p->reclaim->get_device_instance(), op,
                       keys, maxK, klens, values, maxV, vlens, out_values, out_vlens, results, n);
    if (hipGetLastError() != hipSuccess) return 0xFFFFFFFFu;
    if (op != 2) p->tree->drain(0);
  }
  return p->tree->last_error();
}

}  // extern "C"
