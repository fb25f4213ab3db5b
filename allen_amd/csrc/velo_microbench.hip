// Measured peaks of the device the path runs on, for the roofline figures (no reference
// counterpart): vector-ALU FMA issue, LDS read bandwidth and HBM copy bandwidth, each timed with
// HIP events on the caller's stream.
//
//   k_valu_peak   16 independent FMA chains per lane (v_fma_f32, packed FP32 disabled)
//   k_lds_peak    conflict-free 128-bit LDS reads (consecutive lanes, consecutive 16 B)
//   k_copy_peak   a grid-stride copy between two 1 GiB buffers, four 16 B loads in flight per lane
#include "velo_amd.h"
#include "velo_common.cuh"
#include "velo_kernels.h"

namespace velo {
namespace {
constexpr unsigned mb_block = 256;
constexpr unsigned mb_chains = 16;
constexpr unsigned mb_lds_vectors = 4; // uint4 read per lane and iteration

__global__ void __launch_bounds__(mb_block) k_valu_peak(float* out, unsigned iters, float m, float c)
{
  float a[mb_chains];
#pragma unroll
  for (unsigned k = 0; k < mb_chains; ++k) a[k] = static_cast<float>(threadIdx.x + k);
  for (unsigned i = 0; i < iters; ++i) {
#pragma unroll
    for (unsigned k = 0; k < mb_chains; ++k) a[k] = __builtin_fmaf(a[k], m, c);
  }
  float s = 0.f;
#pragma unroll
  for (unsigned k = 0; k < mb_chains; ++k) s += a[k];
  // A store that practically never happens keeps every chain live.
  if (s == 1234.5678f) out[blockIdx.x * mb_block + threadIdx.x] = s;
}

__global__ void __launch_bounds__(mb_block) k_lds_peak(uint32_t* out, unsigned iters)
{
  __shared__ uint4 buf[mb_lds_vectors * mb_block];
#pragma unroll
  for (unsigned k = 0; k < mb_lds_vectors; ++k)
    buf[k * mb_block + threadIdx.x] = make_uint4(threadIdx.x, k, ~threadIdx.x, k ^ 7u);
  __syncthreads();
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  unsigned off = threadIdx.x;
  for (unsigned i = 0; i < iters; ++i) {
    // An empty compiler barrier: the reads are not hoisted out of the loop.
    asm volatile("" ::: "memory");
#pragma unroll
    for (unsigned k = 0; k < mb_lds_vectors; ++k) {
      const uint4 v = buf[k * mb_block + off];
      acc.x ^= v.x;
      acc.y ^= v.y;
      acc.z ^= v.z;
      acc.w ^= v.w;
    }
    off = (off + wave_size) & (mb_block - 1);
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) out[blockIdx.x * mb_block + threadIdx.x] = acc.x;
}

// Four independent 16 B loads per lane and round, then their stores (n is a multiple of
// 4 x gridDim.x x mb_block).
__global__ void __launch_bounds__(mb_block) k_copy_peak(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
  const size_t stride = static_cast<size_t>(gridDim.x) * mb_block;
  for (size_t i = static_cast<size_t>(blockIdx.x) * mb_block + threadIdx.x; i < n; i += 4 * stride) {
    uint4 v[4];
#pragma unroll
    for (unsigned k = 0; k < 4; ++k) v[k] = src[i + k * stride];
#pragma unroll
    for (unsigned k = 0; k < 4; ++k) dst[i + k * stride] = v[k];
  }
}

// Seconds of `reps` launches of `launch` after one untimed launch, HIP events on `s`.
template<typename F>
hipError_t time_launches(F&& launch, unsigned reps, hipStream_t s, double* seconds)
{
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(a, s);
  for (unsigned r = 0; e == hipSuccess && r < reps; ++r) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(b, s);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  *seconds = ms / 1e3;
  if (a) (void) hipEventDestroy(a);
  if (b) (void) hipEventDestroy(b);
  return e;
}
} // namespace
} // namespace velo

namespace velo {
hipError_t run_microbench(int kind, double* result, hipStream_t stream)
{
  *result = 0.0;
  constexpr unsigned reps = 5;
  double seconds = 0.0;
  hipError_t e = hipSuccess;
  if (kind == 0) {
    constexpr unsigned grid = 8192, iters = 2048;
    float* out = nullptr;
    e = hipMalloc(&out, sizeof(float) * grid * mb_block);
    if (e == hipSuccess)
      e = time_launches([&] { hipLaunchKernelGGL(k_valu_peak, dim3(grid), dim3(mb_block), 0, stream, out, iters, 0.999f, 0.001f); },
                        reps, stream, &seconds);
    if (out) (void) hipFree(out);
    if (e == hipSuccess && seconds > 0)
      *result = static_cast<double>(reps) * grid * mb_block * iters * mb_chains / seconds;
  }
  else if (kind == 1) {
    constexpr unsigned grid = 8192, iters = 4096;
    uint32_t* out = nullptr;
    e = hipMalloc(&out, sizeof(uint32_t) * grid * mb_block);
    if (e == hipSuccess)
      e = time_launches([&] { hipLaunchKernelGGL(k_lds_peak, dim3(grid), dim3(mb_block), 0, stream, out, iters); },
                        reps, stream, &seconds);
    if (out) (void) hipFree(out);
    if (e == hipSuccess && seconds > 0)
      *result = static_cast<double>(reps) * grid * mb_block * iters * mb_lds_vectors * sizeof(uint4) / seconds;
  }
  else {
    constexpr size_t bytes = size_t(1) << 30;
    constexpr size_t n = bytes / sizeof(uint4);
    constexpr unsigned grid = 256 * 8;
    static_assert(n % (4 * size_t(grid) * mb_block) == 0, "whole rounds");
    uint4 *src = nullptr, *dst = nullptr;
    e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, stream);
    if (e == hipSuccess)
      e = time_launches([&] { hipLaunchKernelGGL(k_copy_peak, dim3(grid), dim3(mb_block), 0, stream, src, dst, n); },
                        reps, stream, &seconds);
    if (src) (void) hipFree(src);
    if (dst) (void) hipFree(dst);
    if (e == hipSuccess && seconds > 0) *result = static_cast<double>(reps) * 2.0 * bytes / seconds;
  }
  return e;
}
} // namespace velo
