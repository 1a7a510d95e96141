/*
 * The learned cost model's network on the GPU: an MLP 24 -> 64 -> 64 -> 64 -> 1
 * (ReLU hidden layers, FP32) with hand-written forward, backward and Adam
 * kernels.  Sizes are tiny, so every kernel keeps the layer's weights in LDS
 * and gives one thread one output element; reductions over the batch run in
 * a fixed order, so training is deterministic for a given seed.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adatm_kern.h"

#define MB 256

namespace {

/* Z = H W + b (H: B x I, W: I x O row-major), Hout = relu(Z) or Z */
__global__ __launch_bounds__(MB) void k_dense_fwd(float const *__restrict__ H,
                                                  float const *__restrict__ W,
                                                  float const *__restrict__ bias,
                                                  float *__restrict__ Z,
                                                  float *__restrict__ Hout,
                                                  uint32_t B, uint32_t I,
                                                  uint32_t O, int relu)
{
  HIP_DYNAMIC_SHARED(float, sw) /* I*O + O */
  for (uint32_t x = threadIdx.x; x < I * O; x += MB) sw[x] = W[x];
  for (uint32_t x = threadIdx.x; x < O; x += MB) sw[I * O + x] = bias[x];
  __syncthreads();
  uint64_t const e = (uint64_t)blockIdx.x * MB + threadIdx.x;
  if (e >= (uint64_t)B * O) return;
  uint32_t const r = (uint32_t)(e / O), o = (uint32_t)(e - (uint64_t)r * O);
  float s = sw[I * O + o];
  float const *h = H + (size_t)r * I;
  for (uint32_t i = 0; i < I; ++i) s += h[i] * sw[i * O + o];
  Z[e] = s;
  Hout[e] = (relu && s < 0.f) ? 0.f : s;
}

/* dy = 2 (y - t) / B and per-block partial sums of (y - t)^2 */
__global__ __launch_bounds__(MB) void k_mse(float const *__restrict__ y,
                                            float const *__restrict__ t,
                                            uint32_t B, float *__restrict__ dy,
                                            float *__restrict__ part)
{
  __shared__ float red[MB];
  uint32_t const i = blockIdx.x * MB + threadIdx.x;
  float d = 0.f;
  if (i < B) {
    d = y[i] - t[i];
    dy[i] = 2.f * d / (float)B;
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (uint32_t k = MB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* dH = dZ W^T; with Zprev, dZprev = dH * (Zprev > 0) */
__global__ __launch_bounds__(MB) void k_dense_bwd_in(float const *__restrict__ dZ,
                                                     float const *__restrict__ W,
                                                     float const *__restrict__ Zprev,
                                                     float *__restrict__ dOut,
                                                     uint32_t B, uint32_t I,
                                                     uint32_t O)
{
  HIP_DYNAMIC_SHARED(float, sw)
  for (uint32_t x = threadIdx.x; x < I * O; x += MB) sw[x] = W[x];
  __syncthreads();
  uint64_t const e = (uint64_t)blockIdx.x * MB + threadIdx.x;
  if (e >= (uint64_t)B * I) return;
  uint32_t const r = (uint32_t)(e / I), i = (uint32_t)(e - (uint64_t)r * I);
  float s = 0.f;
  float const *dz = dZ + (size_t)r * O;
  for (uint32_t o = 0; o < O; ++o) s += dz[o] * sw[i * O + o];
  if (Zprev && Zprev[e] <= 0.f) s = 0.f;
  dOut[e] = s;
}

/* dW = H^T dZ (I x O), db = sum_b dZ; one thread per weight, batch order */
__global__ __launch_bounds__(MB) void k_dense_bwd_w(float const *__restrict__ H,
                                                    float const *__restrict__ dZ,
                                                    float *__restrict__ dW,
                                                    float *__restrict__ db,
                                                    uint32_t B, uint32_t I,
                                                    uint32_t O)
{
  uint32_t const e = blockIdx.x * MB + threadIdx.x;
  if (e >= (I + 1) * O) return;
  uint32_t const i = e / O, o = e - i * O;
  float s = 0.f;
  if (i < I)
    for (uint32_t r = 0; r < B; ++r) s += H[(size_t)r * I + i] * dZ[(size_t)r * O + o];
  else
    for (uint32_t r = 0; r < B; ++r) s += dZ[(size_t)r * O + o];
  if (i < I)
    dW[e] = s;
  else
    db[o] = s;
}

__global__ void k_adam(float *__restrict__ p, float const *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, uint32_t n,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
  uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float const gi = g[i];
  float const mi = b1 * m[i] + (1.f - b1) * gi;
  float const vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + MB - 1) / MB); }

} // namespace

extern "C" {

int adatm_k_dense_fwd(hipStream_t s, float const *H, float const *W,
                      float const *b, float *Z, float *Hout, uint32_t B,
                      uint32_t I, uint32_t O, int relu)
{
  size_t const lds = ((size_t)I * O + O) * sizeof(float);
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)B * O)), dim3(MB), lds, s,
                     H, W, b, Z, Hout, B, I, O, relu);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int adatm_k_mse(hipStream_t s, float const *y, float const *t, uint32_t B,
                float *dy, float *part)
{
  hipLaunchKernelGGL(k_mse, dim3(blocks(B)), dim3(MB), 0, s, y, t, B, dy, part);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int adatm_k_dense_bwd(hipStream_t s, float const *H, float const *dZ,
                      float const *W, float const *Zprev, float *dH, float *dW,
                      float *db, uint32_t B, uint32_t I, uint32_t O)
{
  if (dH) {
    size_t const lds = (size_t)I * O * sizeof(float);
    hipLaunchKernelGGL(k_dense_bwd_in, dim3(blocks((uint64_t)B * I)), dim3(MB), lds,
                       s, dZ, W, Zprev, dH, B, I, O);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)(I + 1) * O)), dim3(MB),
                     0, s, H, dZ, dW, db, B, I, O);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int adatm_k_adam(hipStream_t s, float *p, float const *g, float *m, float *v,
                 uint32_t n, float lr, float b1, float b2, float eps, float c1,
                 float c2)
{
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(MB), 0, s, p, g, m, v, n, lr,
                     b1, b2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

} /* extern "C" */
