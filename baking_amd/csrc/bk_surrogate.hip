// Return-map surrogate: BF16 matrix-core inference over feature rows and the
// FP32 training kernels (forward, backward, weight gradients, Adam).
//
// Training runs one thread per sample for the forward and backward sweeps
// (the weights are wave-uniform, so they arrive through scalar loads) and
// one thread per parameter for the weight gradients, summed over fixed
// chunks of the batch and then over the chunks in a fixed order, so a
// gradient is the same from run to run.
#include "bk_surrogate.h"
#include "../../include/baking_abi.h"

#include <algorithm>

using namespace bk;

void bk_internal_set_error(const char* what);  // bk_last_error's text (bk_substep.hip)

namespace {
int fail(const char* what) {
  bk_internal_set_error(what);
  return -1;
}
int hip_fail(hipError_t e) {
  bk_internal_set_error(hipGetErrorString(e));
  return -2;
}
inline int div_up(long long a, long long b) { return int((a + b - 1) / b); }
constexpr int kGradChunk = 1024;  // batch rows per partial weight gradient
}  // namespace

__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------- inference
__global__ __launch_bounds__(256) void k_mlp_infer(const uint8_t* __restrict__ blob, const float* __restrict__ rows,
                                                   float* __restrict__ out, int n) {
  __shared__ __align__(16) uint8_t lds[kMlpLdsBytes];
  mlp_rows(blob, rows, out, n, lds);
}

// ---------------------------------------------------------------- FP32 training
// Forward of normalised rows: hidden pre-activations z[l][b][64] (l = 0..3) and normalised outputs y[b][4].
__global__ __launch_bounds__(64) void k_train_forward(const float* __restrict__ th, const float* __restrict__ x, int B,
                                                      float* __restrict__ z, float* __restrict__ y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float a[kMlpWidth], t[kMlpWidth];
  float xin[kMlpRow];
#pragma unroll
  for (int k = 0; k < kMlpRow; ++k) xin[k] = k < kMlpIn ? x[size_t(b) * kMlpRow + k] : 0.0f;
#pragma unroll
  for (int j = 0; j < kMlpWidth; ++j) {
    float s = th[kThetaB1 + j];
#pragma unroll
    for (int k = 0; k < kMlpRow; ++k) s += th[kThetaW1 + j * kMlpRow + k] * xin[k];
    t[j] = s;
  }
#pragma unroll
  for (int j = 0; j < kMlpWidth; ++j) {
    z[size_t(b) * kMlpWidth + j] = t[j];
    a[j] = silu(t[j]);
  }
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const float* W = th + kThetaHidden + l * kThetaHiddenStride;
    const float* bias = W + kMlpWidth * kMlpWidth;
#pragma unroll
    for (int j = 0; j < kMlpWidth; ++j) {
      float s = bias[j];
#pragma unroll
      for (int k = 0; k < kMlpWidth; ++k) s += W[j * kMlpWidth + k] * a[k];
      t[j] = s;
    }
    float* zl = z + (size_t(l + 1) * B + b) * kMlpWidth;
#pragma unroll
    for (int j = 0; j < kMlpWidth; ++j) {
      zl[j] = t[j];
      a[j] = silu(t[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < kMlpOut; ++j) {
    float s = th[kThetaB5 + j];
#pragma unroll
    for (int k = 0; k < kMlpWidth; ++k) s += th[kThetaW5 + j * kMlpWidth + k] * a[k];
    y[size_t(b) * kMlpOut + j] = s;
  }
}

// Backward of the mean squared error over the batch and the 4 outputs: deltas
// d[l][b][64] of the hidden layers (l = 0..3) and d_out[b][4]; loss summed in `loss`.
__global__ __launch_bounds__(64) void k_train_backward(const float* __restrict__ th, const float* __restrict__ z,
                                                       const float* __restrict__ y, const float* __restrict__ target, int B,
                                                       float* __restrict__ d, float* __restrict__ d_out, double* loss) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (b < B) {
    const float inv = 2.0f / float(B * kMlpOut);
    float g[kMlpOut];
#pragma unroll
    for (int j = 0; j < kMlpOut; ++j) {
      const float e = y[size_t(b) * kMlpOut + j] - target[size_t(b) * kMlpOut + j];
      sq += e * e;
      g[j] = inv * e;
      d_out[size_t(b) * kMlpOut + j] = g[j];
    }
    // deltas propagate row by row: nx[k] += W[j][k] * delta[j] streams each weight row once
    float dl[kMlpWidth], nx[kMlpWidth];
    const float* z4 = z + (size_t(3) * B + b) * kMlpWidth;
#pragma unroll
    for (int k = 0; k < kMlpWidth; ++k) nx[k] = 0.0f;
#pragma unroll
    for (int j = 0; j < kMlpOut; ++j)
#pragma unroll
      for (int k = 0; k < kMlpWidth; ++k) nx[k] += th[kThetaW5 + j * kMlpWidth + k] * g[j];
    float* d4 = d + (size_t(3) * B + b) * kMlpWidth;
#pragma unroll
    for (int k = 0; k < kMlpWidth; ++k) {
      dl[k] = nx[k] * silu_grad(z4[k]);
      d4[k] = dl[k];
    }
#pragma unroll
    for (int l = 2; l >= 0; --l) {
      // delta of hidden layer l from that of layer l + 1 through hidden matrix l
      const float* W = th + kThetaHidden + l * kThetaHiddenStride;
      const float* zl = z + (size_t(l) * B + b) * kMlpWidth;
#pragma unroll
      for (int k = 0; k < kMlpWidth; ++k) nx[k] = 0.0f;
#pragma unroll
      for (int j = 0; j < kMlpWidth; ++j)
#pragma unroll
        for (int k = 0; k < kMlpWidth; ++k) nx[k] += W[j * kMlpWidth + k] * dl[j];
      float* dd = d + (size_t(l) * B + b) * kMlpWidth;
#pragma unroll
      for (int k = 0; k < kMlpWidth; ++k) {
        dl[k] = nx[k] * silu_grad(zl[k]);
        dd[k] = dl[k];
      }
    }
  }
  // loss: one atomic per wavefront
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  if (__lane_id() == 0 && sq != 0.0f) atomicAdd(loss, double(sq) / double(B * kMlpOut));
}

// Partial weight gradients over rows [c*kGradChunk, (c+1)*kGradChunk): one thread per parameter.
__global__ __launch_bounds__(64) void k_train_wgrad(const float* __restrict__ x, const float* __restrict__ z,
                                                    const float* __restrict__ d, const float* __restrict__ d_out, int B,
                                                    float* __restrict__ partial) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (p >= kThetaCount) return;
  const int b0 = c * kGradChunk, b1 = min(B, b0 + kGradChunk);
  float s = 0.0f;
  if (p < kThetaB1) {  // W1[j][k]: delta of hidden layer 0 times input k
    const int j = (p - kThetaW1) / kMlpRow, k = (p - kThetaW1) % kMlpRow;
    if (k < kMlpIn)
      for (int b = b0; b < b1; ++b) s += d[size_t(b) * kMlpWidth + j] * x[size_t(b) * kMlpRow + k];
  } else if (p < kThetaHidden) {  // b1[j]
    const int j = p - kThetaB1;
    for (int b = b0; b < b1; ++b) s += d[size_t(b) * kMlpWidth + j];
  } else if (p < kThetaW5) {  // hidden matrix l: delta of layer l + 1 times activation of layer l
    const int l = (p - kThetaHidden) / kThetaHiddenStride, r = (p - kThetaHidden) % kThetaHiddenStride;
    const float* dl = d + size_t(l + 1) * B * kMlpWidth;
    if (r < kMlpWidth * kMlpWidth) {
      const int j = r / kMlpWidth, k = r % kMlpWidth;
      const float* zl = z + size_t(l) * B * kMlpWidth;
      for (int b = b0; b < b1; ++b) s += dl[size_t(b) * kMlpWidth + j] * silu(zl[size_t(b) * kMlpWidth + k]);
    } else {
      const int j = r - kMlpWidth * kMlpWidth;
      for (int b = b0; b < b1; ++b) s += dl[size_t(b) * kMlpWidth + j];
    }
  } else if (p < kThetaB5) {  // W5[j][k]
    const int j = (p - kThetaW5) / kMlpWidth, k = (p - kThetaW5) % kMlpWidth;
    const float* z4 = z + size_t(3) * B * kMlpWidth;
    for (int b = b0; b < b1; ++b) s += d_out[size_t(b) * kMlpOut + j] * silu(z4[size_t(b) * kMlpWidth + k]);
  } else {  // b5[j]
    const int j = p - kThetaB5;
    for (int b = b0; b < b1; ++b) s += d_out[size_t(b) * kMlpOut + j];
  }
  partial[size_t(c) * kThetaCount + p] = s;
}

__global__ void k_sum_chunks(const float* __restrict__ partial, int chunks, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kThetaCount) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += partial[size_t(c) * kThetaCount + p];
  grad[p] = s;
}

__global__ void k_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, int n,
                       float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------- C ABI
extern "C" {

int32_t bk_mlp_theta_count(void) { return kThetaCount; }
int32_t bk_mlp_blob_bytes(void) { return kBlobBytes; }
int32_t bk_mlp_grad_chunk(void) { return kGradChunk; }

int bk_mlp_infer(const void* blob, const float* rows, float* out, int32_t n, void* stream) {
  if (!blob || n < 0) return fail("bk_mlp_infer: bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(div_up(n, 256), 1024)), dim3(256), 0, (hipStream_t)stream, (const uint8_t*)blob,
                     rows, out, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e);
}

int bk_mlp_forward(const float* theta, const float* rows, float* out, int32_t n, float* z, void* stream) {
  if (!theta || n < 0) return fail("bk_mlp_forward: bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_train_forward, dim3(div_up(n, 64)), dim3(64), 0, (hipStream_t)stream, theta, rows, n, z, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e);
}

int bk_mlp_gradient(const float* theta, const float* rows, const float* targets, int32_t n, float* z, float* y, float* d,
                    float* d_out, float* partial, float* grad, double* loss, void* stream) {
  if (!theta || n < 1) return fail("bk_mlp_gradient: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int chunks = div_up(n, kGradChunk);
  hipLaunchKernelGGL(k_train_forward, dim3(div_up(n, 64)), dim3(64), 0, st, theta, rows, n, z, y);
  hipLaunchKernelGGL(k_train_backward, dim3(div_up(n, 64)), dim3(64), 0, st, theta, z, y, targets, n, d, d_out, loss);
  hipLaunchKernelGGL(k_train_wgrad, dim3(div_up(kThetaCount, 64), chunks), dim3(64), 0, st, rows, z, d, d_out, n, partial);
  hipLaunchKernelGGL(k_sum_chunks, dim3(div_up(kThetaCount, 256)), dim3(256), 0, st, partial, chunks, grad);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e);
}

int bk_adam_step(float* theta, const float* grad, float* m, float* v, int32_t count, float lr, float beta1, float beta2, float eps,
                 int32_t step, void* stream) {
  if (!theta || count < 1 || step < 1) return fail("bk_adam_step: bad argument");
  const float bc1 = 1.0f - powf(beta1, float(step)), bc2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_adam, dim3(div_up(count, 256)), dim3(256), 0, (hipStream_t)stream, theta, grad, m, v, count, lr, beta1, beta2,
                     eps, bc1, bc2);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e);
}

}  // extern "C"
