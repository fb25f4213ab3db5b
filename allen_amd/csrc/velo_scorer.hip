// The learned seed scorer: a 10 -> 32 -> 32 -> 1 MLP with ReLU on BF16 MFMA (FP32 accumulation),
// trained on the seeds the plain search produces (no reference counterpart).
//
//   k_mlp_train        forward, binary cross-entropy and backward of 64 seeds per workgroup
//                      (32 per wave), gradients summed per workgroup into partial[block][1441]
//   k_mlp_reduce_adam  sums the partial gradients in block order (deterministic) and applies
//                      one Adam step to the FP32 master parameters
//   k_mlp_score        forward only: the logit of each seed
//
// Arithmetic (the CPU restatement in oracle/mlp.py follows it step for step):
//   xb  = bf16(x)                      Z1 = xb W1b + b1 (FP32 accumulation)     W1b = bf16(W1)
//   A1b = bf16(relu(Z1))               Z2 = A1b W2b + b2                        W2b = bf16(W2)
//   A2  = relu(Z2)                     z3 = A2 . w3b + b3 (FP32)                w3b = bf16(w3)
//   dz3 = (sigmoid(z3) - y) / N        dZ2 = dz3 w3b * [Z2 > 0],  dZ2b = bf16(dZ2)
//   dW2 = A1b^T dZ2b, db2 = sum dZ2    dA1 = dZ2b W2b^T, dZ1 = dA1 * [Z1 > 0], dZ1b = bf16(dZ1)
//   dW1 = xb^T dZ1b, db1 = sum dZ1     dw3 = sum dz3 A2, db3 = sum dz3
// bf16() rounds FP32 to nearest even. Each 32 x 32 product is a chain of
// v_mfma_f32_32x32x8_bf16 over K in steps of 8 with operands read from LDS.
#include "velo_common.cuh"
#include "velo_kernels.h"
#include "velo_mfma.cuh"

namespace velo {
namespace mlp {

// A wave owns 32 rows (the M of one 32 x 32 product), so a workgroup of rows_per_block rows has
// `waves` waves of 64 lanes.
constexpr unsigned rows_per_wave = 32, waves = rows_per_block / rows_per_wave, block_threads = waves * wave_size;

// Shared per workgroup: the bf16 weights (W1 padded to 32 input rows), biases and w3.
struct Weights {
  uint16_t w1[32 * 32]; // W1b[i][j], rows i >= 10 zero
  uint16_t w2[32 * 32]; // W2b[j][k]
  float b1[32], b2[32], w3[32];
  float b3;
};

// Per wave: the activations of its 32 rows ([Z1 > 0] is [a1b != 0]: rounding a positive FP32 to
// bf16 never gives zero).
struct WaveTiles {
  uint16_t xb[32 * 32];  // bf16(x), columns >= 10 zero
  uint16_t a1b[32 * 32];
  float z2[32 * 32];
  uint16_t dz2b[32 * 32];
  uint16_t dz1b[32 * 32];
  float dz3[32];
};

__device__ void load_weights(const float* p, Weights& w)
{
  for (unsigned t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
    const unsigned i = t / 32, j = t % 32;
    w.w1[t] = i < n_inputs ? to_bf16(p[param_w1 + i * 32 + j]) : 0;
    w.w2[t] = to_bf16(p[param_w2 + t]);
  }
  for (unsigned t = threadIdx.x; t < 32; t += blockDim.x) {
    w.b1[t] = p[param_b1 + t];
    w.b2[t] = p[param_b2 + t];
    w.w3[t] = from_bf16(to_bf16(p[param_w3 + t]));
  }
  if (threadIdx.x == 0) w.b3 = p[param_b3];
}

// Forward of the wave's rows (row0 = first row, n rows valid): fills xb, a1b, z2 and
// returns this lane's logit for row l % 32 (both half-waves hold it).
__device__ float forward(const float* x, unsigned row0, unsigned n, const Weights& w, WaveTiles& t)
{
  const unsigned l = lane_id();
  for (unsigned e = l; e < 32 * 32; e += wave_size) {
    const unsigned r = e / 32, c = e % 32;
    t.xb[e] = (r < n && c < n_inputs) ? to_bf16(x[(size_t) (row0 + r) * n_inputs + c]) : 0;
  }
  wave_sync();
  f32x16 acc = mma32(t.xb, 32, 1, w.w1, 32, 1, 32, zero16());
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) {
    const unsigned r = d_row(v), c = d_col();
    const float z = acc[v] + w.b1[c];
    t.a1b[r * 32 + c] = to_bf16(z > 0.f ? z : 0.f);
  }
  wave_sync();
  acc = mma32(t.a1b, 32, 1, w.w2, 32, 1, 32, zero16());
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) {
    const unsigned r = d_row(v), c = d_col();
    t.z2[r * 32 + c] = acc[v] + w.b2[c];
  }
  wave_sync();
  // z3 of row l % 32: columns 16 (l / 32) .. + 16 here, the other half from lane l ^ 32.
  const unsigned r = l % 32, h = l / 32;
  float s = 0.f;
  for (unsigned k = 16 * h; k < 16 * h + 16; ++k) {
    const float z = t.z2[r * 32 + k];
    s += (z > 0.f ? z : 0.f) * w.w3[k];
  }
  const float other = __shfl_xor(s, 32, wave_size);
  const float sum = h == 0 ? s + other : other + s;
  return sum + w.b3;
}

} // namespace mlp

using namespace mlp;

__global__ void __launch_bounds__(block_threads) k_mlp_train(
  const float* features,
  const uint8_t* labels,
  unsigned n,
  float inv_norm,
  const float* params,
  float* partial,
  float* loss_partial)
{
  __shared__ Weights w;
  __shared__ WaveTiles tiles[waves];
  __shared__ float red[waves][n_params];
  const unsigned wave = threadIdx.x / wave_size, l = lane_id();
  WaveTiles& t = tiles[wave];
  load_weights(params, w);
  __syncthreads();
  const unsigned row0 = blockIdx.x * rows_per_block + wave * rows_per_wave;
  const unsigned nv = row0 < n ? min(rows_per_wave, n - row0) : 0u;
  const float z3 = forward(features, row0, nv, w, t);
  const unsigned r = l % 32;
  const bool valid = r < nv;
  const float y = valid ? static_cast<float>(labels[row0 + r]) : 0.f;
  const float p = 1.f / (1.f + expf(-z3));
  const float dz3 = valid ? (p - y) * inv_norm : 0.f;
  // Loss of the row (binary cross-entropy from the logit, stable form), counted once per row.
  const float loss = valid && l < 32 ? (fmaxf(z3, 0.f) - z3 * y + log1pf(expf(-fabsf(z3)))) : 0.f;
  if (l < 32) t.dz3[r] = dz3;
  wave_sync();
  // dZ2 = dz3 w3 [Z2 > 0]
  for (unsigned e = l; e < 32 * 32; e += wave_size) {
    const unsigned rr = e / 32, c = e % 32;
    const float g = t.z2[e] > 0.f ? t.dz3[rr] * w.w3[c] : 0.f;
    t.dz2b[e] = to_bf16(g);
  }
  wave_sync();
  float* g = red[wave];
  // Bias and w3 gradients (FP32, rows summed in order): lane c < 32 owns column c.
  if (l < 32) {
    float db2 = 0.f, dw3 = 0.f;
    for (unsigned rr = 0; rr < 32; ++rr) {
      const float z2 = t.z2[rr * 32 + l];
      const float a2 = z2 > 0.f ? z2 : 0.f;
      db2 += z2 > 0.f ? t.dz3[rr] * w.w3[l] : 0.f;
      dw3 += t.dz3[rr] * a2;
    }
    g[param_b2 + l] = db2;
    g[param_w3 + l] = dw3;
  }
  if (l == 0) {
    float db3 = 0.f;
    for (unsigned rr = 0; rr < 32; ++rr) db3 += t.dz3[rr];
    g[param_b3] = db3;
  }
  // dW2 = A1b^T dZ2b: A(j, r) = a1b[r][j], B(r, k) = dz2b[r][k].
  f32x16 acc = mma32(t.a1b, 1, 32, t.dz2b, 32, 1, 32, zero16());
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) g[param_w2 + d_row(v) * 32 + d_col()] = acc[v];
  // dA1 = dZ2b W2b^T: A(r, k) = dz2b[r][k], B(k, j) = w2[j][k]; dZ1 = dA1 [Z1 > 0].
  acc = mma32(t.dz2b, 32, 1, w.w2, 1, 32, 32, zero16());
  float* dz1 = t.z2; // z2 is no longer needed: reuse it for dZ1 (FP32)
  wave_sync();
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) {
    const unsigned rr = d_row(v), c = d_col();
    const float d = t.a1b[rr * 32 + c] != 0 ? acc[v] : 0.f;
    dz1[rr * 32 + c] = d;
    t.dz1b[rr * 32 + c] = to_bf16(d);
  }
  wave_sync();
  if (l < 32) {
    float db1 = 0.f;
    for (unsigned rr = 0; rr < 32; ++rr) db1 += dz1[rr * 32 + l];
    g[param_b1 + l] = db1;
  }
  // dW1 = xb^T dZ1b: A(i, r) = xb[r][i], B(r, j) = dz1b[r][j]; rows i < 10 are parameters.
  acc = mma32(t.xb, 1, 32, t.dz1b, 32, 1, 32, zero16());
#pragma unroll
  for (unsigned v = 0; v < 16; ++v) {
    const unsigned i = d_row(v);
    if (i < n_inputs) g[param_w1 + i * 32 + d_col()] = acc[v];
  }
  // Loss of the wave's rows.
  float ls = loss;
#pragma unroll
  for (unsigned d = 32; d > 0; d >>= 1) ls += __shfl_xor(ls, d, wave_size);
  __syncthreads();
  // Sum of the waves in wave order.
  for (unsigned q = threadIdx.x; q < n_params; q += blockDim.x) {
    float s = red[0][q];
    for (unsigned k = 1; k < waves; ++k) s += red[k][q];
    partial[(size_t) blockIdx.x * n_params + q] = s;
  }
  __shared__ float s_loss[waves];
  if (l == 0) s_loss[wave] = ls;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (unsigned k = 0; k < waves; ++k) s += s_loss[k];
    loss_partial[blockIdx.x] = s;
  }
}

// One thread per parameter: the gradient is the sum of the blocks' partial gradients in block
// order; Adam with bias correction (step = 1-based count of this update; lr = 0: gradient only).
// Thread 0 also sums the blocks' losses into *loss_out.
__global__ void k_mlp_reduce_adam(
  const float* partial,
  const float* loss_partial,
  unsigned n_blocks,
  float* params,
  float* m,
  float* v,
  float* grad_out,
  float* loss_out,
  float lr,
  float beta1,
  float beta2,
  float eps,
  unsigned step)
{
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0 && loss_out) {
    float s = 0.f;
    for (unsigned b = 0; b < n_blocks; ++b) s += loss_partial[b];
    *loss_out = s;
  }
  if (q >= n_params) return;
  float gsum = 0.f;
  for (unsigned b = 0; b < n_blocks; ++b) gsum += partial[(size_t) b * n_params + q];
  if (grad_out) grad_out[q] = gsum;
  if (lr == 0.f) return;
  const float mq = beta1 * m[q] + (1.f - beta1) * gsum;
  const float vq = beta2 * v[q] + (1.f - beta2) * gsum * gsum;
  m[q] = mq;
  v[q] = vq;
  const float mh = mq / (1.f - powf(beta1, static_cast<float>(step)));
  const float vh = vq / (1.f - powf(beta2, static_cast<float>(step)));
  params[q] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void __launch_bounds__(block_threads) k_mlp_score(
  const float* features,
  unsigned n,
  const float* params,
  float* logits)
{
  __shared__ Weights w;
  __shared__ WaveTiles tiles[waves];
  const unsigned wave = threadIdx.x / wave_size, l = lane_id();
  load_weights(params, w);
  __syncthreads();
  const unsigned row0 = blockIdx.x * rows_per_block + wave * rows_per_wave;
  const unsigned nv = row0 < n ? min(rows_per_wave, n - row0) : 0u;
  const float z3 = forward(features, row0, nv, w, tiles[wave]);
  if (l < 32 && l < nv) logits[row0 + l] = z3;
}

hipError_t launch_mlp_train(
  const float* features,
  const uint8_t* labels,
  unsigned n,
  float inv_norm,
  const float* params,
  float* partial,
  float* loss_partial,
  hipStream_t s)
{
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(
    k_mlp_train, dim3(mlp_blocks(n)), dim3(block_threads), 0, s, features, labels, n, inv_norm, params, partial,
    loss_partial);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce_adam(
  const float* partial,
  const float* loss_partial,
  unsigned n_blocks,
  float* params,
  float* m,
  float* v,
  float* grad_out,
  float* loss_out,
  float lr,
  unsigned step,
  hipStream_t s)
{
  hipLaunchKernelGGL(
    k_mlp_reduce_adam, dim3((n_params + 255) / 256), dim3(256), 0, s, partial, loss_partial, n_blocks, params, m, v,
    grad_out, loss_out, lr, 0.9f, 0.999f, 1e-8f, step);
  return hipGetLastError();
}

hipError_t launch_mlp_score(const float* features, unsigned n, const float* params, float* logits, hipStream_t s)
{
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_score, dim3(mlp_blocks(n)), dim3(block_threads), 0, s, features, n, params, logits);
  return hipGetLastError();
}

} // namespace velo
