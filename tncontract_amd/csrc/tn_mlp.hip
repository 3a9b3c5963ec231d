// Rank-prediction MLP (34 -> 64 -> 64 -> 1, ReLU, FP32): forward, backward
// and Adam, written for the learned-rank truncation (no reference
// counterpart: the reference always runs a full SVD, tensor.py:915).
//
// Parameter vector layout (P = 6465 floats):
//   W1 [34][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3 [1]
// One 64-thread block per sample: thread j owns hidden unit j of both layers.
// Gradients are written per sample (deterministic) and reduced by a second
// kernel, so training is bit-reproducible.
#include "tn_common.h"

namespace tn {

constexpr int MI = 34, MH = 64;
constexpr int OW1 = 0, OB1 = OW1 + MI * MH, OW2 = OB1 + MH, OB2 = OW2 + MH * MH, OW3 = OB2 + MH,
              OB3 = OW3 + MH, MP = OB3 + 1;

__global__ void __launch_bounds__(64) mlp_fwd_kernel(const float* __restrict__ X, int B,
                                                     const float* __restrict__ P,
                                                     float* __restrict__ Y,
                                                     float* __restrict__ H1,
                                                     float* __restrict__ H2) {
  __shared__ float x[MI], h1[MH];
  const int s = blockIdx.x, j = threadIdx.x;
  if (s >= B) return;
  if (j < MI) x[j] = X[(int64_t)s * MI + j];
  __syncthreads();
  float a = P[OB1 + j];
  for (int i = 0; i < MI; ++i) a = fmaf(x[i], P[OW1 + i * MH + j], a);
  a = fmaxf(a, 0.f);
  h1[j] = a;
  if (H1) H1[(int64_t)s * MH + j] = a;
  __syncthreads();
  float b = P[OB2 + j];
  for (int k = 0; k < MH; ++k) b = fmaf(h1[k], P[OW2 + k * MH + j], b);
  b = fmaxf(b, 0.f);
  if (H2) H2[(int64_t)s * MH + j] = b;
  float y = b * P[OW3 + j];
  y = wave_sum(y);
  if (j == 0) Y[s] = y + P[OB3];
}

// Per-sample gradient of L = (1/B) sum_s (y_s - t_s)^2 into G[s][P].
__global__ void __launch_bounds__(64) mlp_bwd_kernel(const float* __restrict__ X,
                                                     const float* __restrict__ T, int B,
                                                     const float* __restrict__ P,
                                                     const float* __restrict__ Y,
                                                     const float* __restrict__ H1,
                                                     const float* __restrict__ H2,
                                                     float* __restrict__ G) {
  __shared__ float x[MI], h1[MH], d2[MH], d1[MH];
  const int s = blockIdx.x, j = threadIdx.x;
  if (s >= B) return;
  float* g = G + (int64_t)s * MP;
  if (j < MI) x[j] = X[(int64_t)s * MI + j];
  h1[j] = H1[(int64_t)s * MH + j];
  const float h2 = H2[(int64_t)s * MH + j];
  const float dy = 2.f * (Y[s] - T[s]) / (float)B;
  __syncthreads();
  g[OW3 + j] = dy * h2;
  if (j == 0) g[OB3] = dy;
  const float dh2 = h2 > 0.f ? dy * P[OW3 + j] : 0.f;
  d2[j] = dh2;
  g[OB2 + j] = dh2;
  __syncthreads();
  for (int k = 0; k < MH; ++k) g[OW2 + k * MH + j] = h1[k] * d2[j];
  float a = 0.f;
  for (int q = 0; q < MH; ++q) a = fmaf(P[OW2 + j * MH + q], d2[q], a);
  const float dh1 = h1[j] > 0.f ? a : 0.f;
  d1[j] = dh1;
  g[OB1 + j] = dh1;
  __syncthreads();
  for (int i = 0; i < MI; ++i) g[OW1 + i * MH + j] = x[i] * d1[j];
}

__global__ void mlp_reduce_kernel(const float* __restrict__ G, int B, float* __restrict__ grad) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MP) return;
  float acc = 0.f;
  for (int s = 0; s < B; ++s) acc += G[(int64_t)s * MP + p];
  grad[p] = acc;
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float gi = g[p];
  float mi = b1 * m[p] + (1.f - b1) * gi;
  float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace tn

using namespace tn;

extern "C" {

int tn_mlp_param_count(void) { return MP; }

int tn_mlp_forward(const float* X, int B, const float* P, float* Y, float* H1, float* H2,
                   void* stream) {
  TN_REQUIRE(B >= 0 && X && P && Y, "tn_mlp_forward: bad arguments");
  if (B == 0) return TN_OK;
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(B), dim3(64), 0, as_stream(stream), X, B, P, Y, H1, H2);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_mlp_backward(const float* X, const float* T, int B, const float* P, const float* Y,
                    const float* H1, const float* H2, float* grad, float* per_sample,
                    void* stream) {
  TN_REQUIRE(B > 0 && X && T && P && Y && H1 && H2 && grad && per_sample,
             "tn_mlp_backward: bad arguments");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(mlp_bwd_kernel, dim3(B), dim3(64), 0, st, X, T, B, P, Y, H1, H2, per_sample);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((MP + 255) / 256), dim3(256), 0, st, per_sample, B,
                     grad);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_adam_step(float* P, const float* grad, float* m, float* v, int n, float lr, float beta1,
                 float beta2, float eps, int step, void* stream) {
  TN_REQUIRE(n > 0 && P && grad && m && v && step >= 1, "tn_adam_step: bad arguments");
  float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), P, grad,
                     m, v, n, lr, beta1, beta2, eps, c1, c2);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- Gaussian sketch
// Counter-based normal samples for the randomized range finder: SplitMix64 of
// (seed, index) -> two uniforms -> Box-Muller.  Deterministic for a seed.
namespace tn {
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void randn_kernel(T* __restrict__ out, int64_t n, unsigned long long seed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    unsigned long long h = splitmix64(seed ^ splitmix64((unsigned long long)i));
    double u1 = ((h >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    double u2 = (double)(splitmix64(h) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = (T)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}
}  // namespace tn

extern "C" int tn_randn(void* out, int dtype, int64_t n, unsigned long long seed, void* stream) {
  TN_REQUIRE(n >= 0 && (out || n == 0), "tn_randn: bad arguments");
  if (n == 0) return TN_OK;
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(randn_kernel<double>, dim3(blocks), dim3(256), 0, as_stream(stream),
                       (double*)out, n, seed);
  else
    hipLaunchKernelGGL(randn_kernel<float>, dim3(blocks), dim3(256), 0, as_stream(stream),
                       (float*)out, n, seed);
  TN_CHECK_LAUNCH();
  return TN_OK;
}
