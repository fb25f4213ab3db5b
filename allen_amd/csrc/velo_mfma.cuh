// BF16 matrix-core helpers of the seed scorer (v_mfma_f32_32x32x8_bf16, FP32 accumulation).
#pragma once
#include "velo_common.cuh"

namespace velo {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float f)
{
  const uint32_t b = __float_as_uint(f);
  return static_cast<uint16_t>((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// D (32 x 32) += A (32 x K) B (K x 32), K a multiple of 8; A(r, k) = A[r * ar + k * ak],
// B(k, c) = B[k * bk + c * bc] (bf16 bits in LDS). Lane l supplies A(l % 32, k0 + 4 (l / 32) + i)
// and B(k0 + 4 (l / 32) + i, l % 32), i = 0..3, and receives D(8 (v / 4) + 4 (l / 32) + v % 4,
// l % 32) in element v (the 32x32x8 operand layout; checked on the device by the tests).
__device__ __forceinline__ f32x16 mma32(
  const uint16_t* A,
  unsigned ar,
  unsigned ak,
  const uint16_t* B,
  unsigned bk,
  unsigned bc,
  unsigned K,
  f32x16 acc)
{
  const unsigned l = lane_id();
  const unsigned rc = l % 32, kq = 4 * (l / 32);
  for (unsigned k0 = 0; k0 < K; k0 += 8) {
    bf16x4 a, b;
#pragma unroll
    for (unsigned i = 0; i < 4; ++i) {
      a[i] = static_cast<short>(A[rc * ar + (k0 + kq + i) * ak]);
      b[i] = static_cast<short>(B[(k0 + kq + i) * bk + rc * bc]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ unsigned d_row(unsigned v) { return 8 * (v / 4) + 4 * (lane_id() / 32) + v % 4; }
__device__ __forceinline__ unsigned d_col() { return lane_id() % 32; }

__device__ __forceinline__ f32x16 zero16()
{
  f32x16 z;
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) z[v] = 0.f;
  return z;
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

} // namespace mlp
} // namespace velo
