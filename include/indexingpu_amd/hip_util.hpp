// HIP helpers shared by the MI355X Masstree headers: error checks and the agent-scope
// atomics used on the pool.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "layout.hpp"

namespace indexingpu_amd {

struct hip_error : std::runtime_error {
  explicit hip_error(const std::string& s) : std::runtime_error(s) {}
};

#define IGPU_HIP_TRY(call)                                                                  \
  do {                                                                                      \
    hipError_t igpu_err_ = (call);                                                          \
    if (igpu_err_ != hipSuccess) {                                                          \
      throw ::indexingpu_amd::hip_error(std::string("HIP error ") +                         \
                                        hipGetErrorString(igpu_err_) + " at " + __FILE__ + \
                                        ":" + std::to_string(__LINE__));                    \
    }                                                                                       \
  } while (0)

namespace dev {

// Coherence scheme of the concurrent paths.  Every access to tree memory that another
// wave may touch during the same launch (node slots, headers, records) is an
// agent-scope atomic: loads and stores carry the sc1 cache policy, so they are served
// coherently across the eight XCDs without whole-cache maintenance.  Ordering between
// these accesses only needs completion, which `order()` provides (s_waitcnt vmcnt(0)
// plus a compiler barrier).  The C++ agent-scope fences would instead write back /
// invalidate the whole L2 on every lock and every validated read (buffer_wbl2 sc1 /
// buffer_inv sc1), which is what non-atomic accesses need and these paths do not have.
__device__ __forceinline__ void order() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ u32 atomic_or(u32* p, u32 v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_and(u32* p, u32 v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_add(u32* p, u32 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 coherent_load(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 coherent_load(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coherent_store(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coherent_store(u32* p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// first error wins
__device__ __forceinline__ void set_error(u32* err, u32 code) {
  u32 expected = 0;
  __hip_atomic_compare_exchange_strong(err, &expected, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev
}  // namespace indexingpu_amd
