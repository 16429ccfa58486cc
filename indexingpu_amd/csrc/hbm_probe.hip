// Random 64-byte line reads over a large device buffer: the measured ceiling of the
// access pattern an index lookup makes (independent 64-B lines at random addresses).
// Four lanes read one line (16 B each, one dwordx4 load); line addresses come from a
// hash of (line stream, iteration), so no index array is read and nothing but the
// lines themselves moves through HBM.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__device__ __forceinline__ uint64_t fmix64(uint64_t v) {
  v ^= v >> 33;
  v *= 0xFF51AFD7ED558CCDull;
  v ^= v >> 33;
  v *= 0xC4CEB9FE1A85EC53ull;
  v ^= v >> 33;
  return v;
}

constexpr uint32_t kThreads = 256;

__global__ void __launch_bounds__(kThreads)
    random_lines_kernel(const uint4* __restrict__ buf, uint64_t line_mask, uint32_t iters,
                        uint32_t* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint32_t quarter = threadIdx.x & 3u;
  const uint64_t stream = t >> 2;
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t k = 0; k < iters; k++) {
    const uint64_t line = fmix64(stream * 0x9E3779B97F4A7C15ull + k) & line_mask;
    const uint4 v = buf[line * 4 + quarter];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  // one vector store per lane keeps the loads live; the buffer holds one word per lane
  sink[t] = acc;
}

}  // namespace

extern "C" {

// buf: device buffer of 2^log2_lines 64-byte lines.  Reads lines_per_stream random
// lines per 4-lane stream with blocks x 256 threads; sink must hold blocks * 256 words.
// Returns 0 or a hipError_t value.
int hbm_random_lines(const void* buf, uint32_t log2_lines, uint32_t blocks,
                     uint32_t lines_per_stream, uint32_t* sink, void* stream) {
  if (!buf || !sink || log2_lines < 1 || log2_lines > 40 || blocks == 0) return (int)hipErrorInvalidValue;
  const uint64_t mask = (1ull << log2_lines) - 1;
  hipLaunchKernelGGL(random_lines_kernel, dim3(blocks), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const uint4*>(buf),
                     mask, lines_per_stream, sink);
  return (int)hipGetLastError();
}

}  // extern "C"
