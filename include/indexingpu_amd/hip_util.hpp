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

__device__ __forceinline__ u32 atomic_or_acq(u32* p, u32 v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_and_rel(u32* p, u32 v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_add_rel(u32* p, u32 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_add_relaxed(u32* p, u32 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_or_relaxed(u32* p, u32 v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 atomic_and_relaxed(u32* p, u32 v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 load_relaxed(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 load_relaxed(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// first error wins
__device__ __forceinline__ void set_error(u32* err, u32 code) {
  u32 expected = 0;
  __hip_atomic_compare_exchange_strong(err, &expected, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev
}  // namespace indexingpu_amd
