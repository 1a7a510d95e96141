/*
 * Dense CP-ALS steps of the sweep on MI355X (gfx950): Gram matrices,
 * (hadamard of Grams)^-1, A = M V^-1 with column normalisation, fit, and the
 * FP64 <-> path-precision conversions.  FP64 arithmetic throughout
 * (cpd_als_iterate_adaptive, src/adatm_cpd.c:276-340, and the SPLATT matrix
 * routines it calls: mat_aTa, calc_gram_inv, mat_matmul, mat_normalize).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adatm_kern.h"
#include "common.hpp"

namespace {

/* ---- misc --------------------------------------------------------------- */
__global__ void k_zero_words(uint32_t *p, size_t nwords)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t const stride = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

__global__ void k_zero_bytes(uint8_t *p, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
__global__ void k_convert_in(double const *__restrict__ src, T *__restrict__ dst,
                             bf16raw *__restrict__ dbf, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    T const v = (T)src[i];
    dst[i] = v;
    if (dbf) dbf[i].v = to_bf16((float)v);
  }
}

template <typename T>
__global__ void k_convert_out(T const *__restrict__ src, double *__restrict__ dst,
                              uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (double)src[i];
}

/* ---- dense CP-ALS steps ------------------------------------------------- */
#define DB 256          /* threads per dense block */
#define NBLK_MAX 256    /* partial-sum blocks */
/* rows per Gram block at least: a block writes an R x R FP64 partial, so
   with fewer rows than ~4R its partial outweighs the rows it reads (the
   batched sweeps' many small factors) */
#define GRAM_ROWS 128u

/* Gram of FP64 rows on the matrix cores: v_mfma_f64_16x16x4f64 takes four
 * rows of A per instruction.  Columns are cut into NT tiles of 16; with
 * a = A[r0 + lane/16][16i + lane%16] and b = A[r0 + lane/16][16j + lane%16]
 * the instruction's A operand (16 x 4, element [m][k] from lane m + 16k) is
 * tile i of the rows transposed and its B operand (4 x 16, [k][n] from lane
 * n + 16k) tile j, so the 16 x 16 accumulator gathers tile (i, j) of A^T A;
 * accumulator register q of lane l holds element (row 4q + l/16, column
 * l%16).  The tiles i <= j of the symmetric result are accumulated (1, 3, 6
 * or 10 for R <= 16, 32, 48, 64), columns >= R are zero.  Each block sums
 * its four waves' tiles in a fixed order and writes one FP64 partial
 * R x R (summed across blocks by k_sum_parts).  With lambda, the rows are
 * first divided by it and written back. */
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int NT>
__device__ __forceinline__ void gram_mfma64_body(uint32_t bid, uint32_t nblk,
                                                 double *__restrict__ A, uint64_t I,
                                                 uint32_t R, double const *__restrict__ lambda,
                                                 double *__restrict__ part)
{
  constexpr int P = NT * (NT + 1) / 2;
  __shared__ double red[3][P][256];
  uint32_t const lane = threadIdx.x & 63u;
  uint32_t const w = threadIdx.x >> 6;
  uint32_t const col = lane & 15u, kq = lane >> 4;
  uint64_t const chunk = ((I + nblk - 1) / nblk + 15) & ~(uint64_t)15;
  uint64_t const lo = (uint64_t)bid * chunk;
  uint64_t const hi = lo + chunk < I ? lo + chunk : I;
  double lam[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
    lam[t] = (lambda && 16u * t + col < R) ? lambda[16u * t + col] : 1.0;
  doublex4 acc[P];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
  /* the wave's row quads rp, rp + 16, ... taken GU at a time: their loads
     are issued together, then their products accumulated in row order */
  constexpr int GU = 2;
  for (uint64_t rp = lo + 4u * w; rp < hi; rp += 16u * GU) {
    double v[GU][NT];
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      uint64_t const r = rp + 16u * (uint32_t)u + kq;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        uint32_t const c = 16u * t + col;
        v[u][t] = (r < hi && c < R) ? A[(size_t)r * R + c] : 0.0;
      }
    }
    if (lambda) {
#pragma unroll
      for (int u = 0; u < GU; ++u) {
        uint64_t const r = rp + 16u * (uint32_t)u + kq;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          uint32_t const c = 16u * t + col;
          if (r < hi && c < R) {
            v[u][t] = v[u][t] / lam[t];
            A[(size_t)r * R + c] = v[u][t];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      if (rp + 16u * (uint32_t)u >= hi) break; /* wave-uniform */
      int p = 0;
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = i; j < NT; ++j, ++p)
          acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u][i], v[u][j], acc[p], 0, 0, 0);
    }
  }
  if (w > 0) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) red[w - 1][p][q * 64 + lane] = acc[p][q];
  }
  __syncthreads();
  if (w != 0) return;
  for (uint32_t ww = 0; ww < 3; ++ww)
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += red[ww][p][q * 64 + lane];
  double *myp = part + (size_t)bid * R * R;
  int p = 0;
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = i; j < NT; ++j, ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        uint32_t const gr = 16u * i + 4u * (uint32_t)q + kq, gc = 16u * j + col;
        if (gr < R && gc < R) {
          myp[(size_t)gr * R + gc] = acc[p][q];
          if (i != j) myp[(size_t)gc * R + gr] = acc[p][q];
        }
      }
}

template <int NT>
__global__ __launch_bounds__(DB) void k_gram_mfma64(double *__restrict__ A, uint64_t I,
                                                    uint32_t R,
                                                    double const *__restrict__ lambda,
                                                    double *__restrict__ part)
{
  gram_mfma64_body<NT>(blockIdx.x, gridDim.x, A, I, R, lambda, part);
}

/* Gram of FP32 rows on the matrix cores: v_mfma_f32_32x32x2f32 takes two
 * rows of A per instruction.  With a = b = A[r0 + lane/32][c0 + lane%32],
 * the instruction's A operand is A^T (32 x 2, element [m][k] from lane
 * m + 32k) and its B operand A (2 x 32), so the 32 x 32 accumulator gathers
 * A^T A over the rows it has seen; accumulator register i of lane l holds
 * element (row (i%4) + 4(l/32) + 8(i/4), column l%32).  R <= 32 uses one
 * tile, R <= 64 the three tiles (0,0), (0,1), (1,1) of the symmetric result;
 * columns >= R are zero.  Each block sums its waves' tiles in a fixed order
 * and writes one FP64 partial R x R (summed across blocks in FP64 by
 * k_sum_parts).  With lambda, the rows are first divided by it and written
 * back (and to the bf16 copy). */
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void gram_mfma_body(uint32_t bid, uint32_t nblk, float *__restrict__ A,
                                                  bf16raw *__restrict__ Abf,
                                                  uint64_t I, uint32_t R,
                                                  double const *__restrict__ lambda,
                                                  double *__restrict__ part)
{
  __shared__ float red[3][3][1024];
  uint32_t const lane = threadIdx.x & 63u;
  uint32_t const w = threadIdx.x >> 6;
  uint32_t const c0 = lane & 31u, c1 = 32u + (lane & 31u);
  bool const two = R > 32;
  uint64_t const chunk = ((I + nblk - 1) / nblk + 7) & ~(uint64_t)7;
  uint64_t const lo = (uint64_t)bid * chunk;
  uint64_t const hi = lo + chunk < I ? lo + chunk : I;
  double const l0 = (lambda && c0 < R) ? lambda[c0] : 1.0;
  double const l1 = (lambda && two && c1 < R) ? lambda[c1] : 1.0;
  floatx16 a00, a01, a11;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    a00[i] = 0.f;
    a01[i] = 0.f;
    a11[i] = 0.f;
  }
  /* the wave's row pairs rp, rp + 8, ... taken GU at a time: their loads
     are issued together, then their products accumulated in row order */
  constexpr int GU = 4;
  for (uint64_t rp = lo + 2u * w; rp < hi; rp += 8u * GU) {
    float v0[GU], v1[GU];
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      uint64_t const r = rp + 8u * (uint32_t)u + (lane >> 5);
      v0[u] = 0.f;
      v1[u] = 0.f;
      if (r < hi) {
        size_t const g0 = (size_t)r * R + c0, g1 = (size_t)r * R + c1;
        if (c0 < R) v0[u] = A[g0];
        if (two && c1 < R) v1[u] = A[g1];
      }
    }
    if (lambda) {
#pragma unroll
      for (int u = 0; u < GU; ++u) {
        uint64_t const r = rp + 8u * (uint32_t)u + (lane >> 5);
        if (r < hi) {
          size_t const g0 = (size_t)r * R + c0, g1 = (size_t)r * R + c1;
          if (c0 < R) {
            v0[u] = (float)((double)v0[u] / l0);
            A[g0] = v0[u];
            if (Abf) Abf[g0].v = to_bf16(v0[u]);
          }
          if (two && c1 < R) {
            v1[u] = (float)((double)v1[u] / l1);
            A[g1] = v1[u];
            if (Abf) Abf[g1].v = to_bf16(v1[u]);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      if (rp + 8u * (uint32_t)u >= hi) break; /* wave-uniform */
      a00 = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[u], v0[u], a00, 0, 0, 0);
      if (two) {
        a01 = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[u], v1[u], a01, 0, 0, 0);
        a11 = __builtin_amdgcn_mfma_f32_32x32x2f32(v1[u], v1[u], a11, 0, 0, 0);
      }
    }
  }
  if (w > 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      red[w - 1][0][i * 64 + lane] = a00[i];
      red[w - 1][1][i * 64 + lane] = a01[i];
      red[w - 1][2][i * 64 + lane] = a11[i];
    }
  }
  __syncthreads();
  if (w != 0) return;
  for (uint32_t ww = 0; ww < 3; ++ww)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      a00[i] += red[ww][0][i * 64 + lane];
      a01[i] += red[ww][1][i * 64 + lane];
      a11[i] += red[ww][2][i * 64 + lane];
    }
  double *myp = part + (size_t)bid * R * R;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    uint32_t const row = (uint32_t)(i % 4) + 4u * (lane >> 5) + 8u * (uint32_t)(i / 4);
    uint32_t const col = lane & 31u;
    if (row < R && col < R) myp[(size_t)row * R + col] = (double)a00[i];
    if (two) {
      if (row < R && 32u + col < R) {
        myp[(size_t)row * R + 32u + col] = (double)a01[i];
        myp[(size_t)(32u + col) * R + row] = (double)a01[i];
      }
      if (32u + row < R && 32u + col < R)
        myp[(size_t)(32u + row) * R + 32u + col] = (double)a11[i];
    }
  }
}

__global__ __launch_bounds__(DB) void k_gram_mfma(float *__restrict__ A,
                                                  bf16raw *__restrict__ Abf,
                                                  uint64_t I, uint32_t R,
                                                  double const *__restrict__ lambda,
                                                  double *__restrict__ part)
{
  gram_mfma_body(blockIdx.x, gridDim.x, A, Abf, I, R, lambda, part);
}

/* out[e] = sum over the nblk partial rows of part[b*len + e] in a fixed
 * order: 8 interleaved chunk sums per element (chunk c takes rows c, c+8,
 * ...), combined in chunk order.  Block: 32 elements x 8 chunks. */
__device__ __forceinline__ void sum_parts_body(uint32_t bid, double const *__restrict__ part,
                                                  uint32_t nblk, uint32_t len,
                                                  double *__restrict__ out)
{
  __shared__ double red[DB];
  uint32_t const c = threadIdx.x >> 5, el = threadIdx.x & 31u;
  uint32_t const e = bid * 32u + el;
  double s = 0.0;
  if (e < len) {
    /* four partial rows' loads in flight, added in row order */
    uint32_t b = c;
    for (; b + 24u < nblk; b += 32u) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = part[(size_t)(b + 8u * u) * len + e];
#pragma unroll
      for (int u = 0; u < 4; ++u) s += x[u];
    }
    for (; b < nblk; b += 8) s += part[(size_t)b * len + e];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (c == 0 && e < len) {
    double t = red[el];
#pragma unroll
    for (uint32_t q = 1; q < 8; ++q) t += red[q * 32u + el];
    out[e] = t;
  }
}

__global__ __launch_bounds__(DB) void k_sum_parts(double const *__restrict__ part,
                                                  uint32_t nblk, uint32_t len,
                                                  double *__restrict__ out)
{
  sum_parts_body(blockIdx.x, part, nblk, len, out);
}

/* Vinv = (hadamard_{k != m} G_k)^-1 by Gauss-Jordan without pivoting (the
 * matrix is symmetric positive definite).  The matrix lives in registers:
 * thread (s, j), s = tid / 64, j = tid % 64, holds column j of rows s, s+4,
 * ..., s+4(Q-1) (R <= 4Q <= 64); each pivot step broadcasts the pivot row and
 * the pivot column through LDS, double-buffered so that one barrier per
 * pivot suffices, and updates branch-free.  One block. */
template <int Q>
__device__ __forceinline__ void solve_inv_body(double const *__restrict__ G,
                                                  uint32_t N, uint32_t m,
                                                  uint32_t R,
                                                  double *__restrict__ Vinv,
                                                  int *__restrict__ status)
{
  __shared__ double rowk[2][64], colk[2][4 * Q];
  uint32_t const s = threadIdx.x >> 6, j = threadIdx.x & 63u;
  double a[Q];
  /* the matrix (plus ridge on the diagonal) into registers.  The Grams are
     read four modes at a time with every load issued before the first
     product (addresses clamped into the buffer, a factor of exactly 1 for
     the skipped mode m and the modes past N), so a build costs ceil(N/4)
     memory round trips rather than one per mode and row; the products are
     taken in mode order as before. */
  auto build = [&](double ridge) {
    uint32_t xo[Q];
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint32_t const i = s + 4u * (uint32_t)q;
      xo[q] = (i < R && j < R) ? i * R + j : 0u;
      v[q] = 1.0;
    }
    for (uint32_t k0 = 0; k0 < N; k0 += 4) {
      double g[4][Q];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint32_t const k = k0 + (uint32_t)u;
        size_t const kb = (size_t)(k < N ? k : 0u) * R * R;
#pragma unroll
        for (int q = 0; q < Q; ++q) g[u][q] = G[kb + xo[q]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint32_t const k = k0 + (uint32_t)u;
        bool const use = k < N && k != m;
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] *= use ? g[u][q] : 1.0;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint32_t const i = s + 4u * (uint32_t)q;
      double x = 1.0;
      if (i < R && j < R) {
        x = v[q];
        if (i == j) x += ridge;
      }
      a[q] = x;
    }
  };
  build(0.0);
  int bad = 0, reg = 0;
  /* a pivot that is not positive (the product is not numerically SPD):
     solve again with a ridge of 1e-10, then 1e-6, of the largest diagonal */
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (attempt > 0) {
      /* the diagonal through LDS, its maximum read by every thread */
      build(0.0);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        uint32_t const i = s + 4u * (uint32_t)q;
        if (i == j) colk[0][i] = a[q];
      }
      __syncthreads();
      double dm = 0.0;
      for (uint32_t i = 0; i < R; ++i) dm = colk[0][i] > dm ? colk[0][i] : dm;
      __syncthreads();
      build(dm * (attempt == 1 ? 1e-10 : 1e-6));
    }
    bad = 0;
    /* entries of rows or columns >= R never flow into rows and columns < R */
    for (uint32_t k = 0; k < R; ++k) {
      uint32_t const b = k & 1u;
      if ((k & 3u) == s) {
        double v = a[0];
#pragma unroll
        for (int q = 1; q < Q; ++q) v = ((uint32_t)q == (k >> 2)) ? a[q] : v;
        rowk[b][j] = v;
      }
      if (j == k) {
#pragma unroll
        for (int q = 0; q < Q; ++q) colk[b][s + 4u * (uint32_t)q] = a[q];
      }
      __syncthreads();
      double const p = rowk[b][k];
      bad |= !(p > 0.0);
      /* an attempt that will be retried stops at its first bad pivot (every
         thread reads the same p, so all leave together); the last attempt
         runs to the end */
      if (bad && attempt < 2) break;
      double const pinv = 1.0 / p;
      double const rk = ((j == k) ? 1.0 : rowk[b][j]) * pinv;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        uint32_t const i = s + 4u * (uint32_t)q;
        double const f = colk[b][i];
        a[q] = (i == k) ? rk : ((j == k) ? 0.0 : a[q]) - f * rk;
      }
    }
    if (!bad) break;
    reg = 1;
    __syncthreads(); /* the LDS broadcasts are reused by the next attempt */
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    uint32_t const i = s + 4u * (uint32_t)q;
    if (i < R && j < R) Vinv[(size_t)i * R + j] = a[q];
  }
  /* 0: clean; 1: solved with a ridge; 2: no positive pivots even then.
     The status is sticky (the worst since it was last cleared): solves run
     in stream order, one writer each, so a plain read-compare-write holds. */
  if (threadIdx.x == 0 && status) {
    int const v = bad ? 2 : reg;
    if (v > *status) *status = v;
  }
}

/* R <= 32: the same elimination in one wave, so a pivot needs no barrier.
 * Lane l = 32h + j holds column j of rows h, h+2, ..., h+30; the pivot row
 * and column go through LDS, whose accesses one wave performs in program
 * order.  Every element is updated by the same expression as in
 * solve_inv_body, from the same operands, so both give identical bits. */
__device__ __forceinline__ void solve_inv32_body(double const *__restrict__ G, uint32_t N,
                                                 uint32_t m, uint32_t R,
                                                 double *__restrict__ Vinv,
                                                 int *__restrict__ status)
{
  constexpr int Q = 16;
  __shared__ double rowk[32], colk[32];
  if (threadIdx.x >= 64u) return;
  uint32_t const h = threadIdx.x >> 5, j = threadIdx.x & 31u;
  double a[Q];
  auto build = [&](double ridge) {
    uint32_t xo[Q];
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint32_t const i = h + 2u * (uint32_t)q;
      xo[q] = (i < R && j < R) ? i * R + j : 0u;
      v[q] = 1.0;
    }
    for (uint32_t k0 = 0; k0 < N; k0 += 4) {
      double g[4][Q];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint32_t const k = k0 + (uint32_t)u;
        size_t const kb = (size_t)(k < N ? k : 0u) * R * R;
#pragma unroll
        for (int q = 0; q < Q; ++q) g[u][q] = G[kb + xo[q]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint32_t const k = k0 + (uint32_t)u;
        bool const use = k < N && k != m;
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] *= use ? g[u][q] : 1.0;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint32_t const i = h + 2u * (uint32_t)q;
      double x = 1.0;
      if (i < R && j < R) {
        x = v[q];
        if (i == j) x += ridge;
      }
      a[q] = x;
    }
  };
  build(0.0);
  int bad = 0, reg = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (attempt > 0) {
      /* the diagonal through LDS, its maximum read by every lane */
      build(0.0);
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (h + 2u * (uint32_t)q == j) colk[j] = a[q];
      __builtin_amdgcn_wave_barrier();
      double dm = 0.0;
      for (uint32_t i = 0; i < R; ++i) dm = colk[i] > dm ? colk[i] : dm;
      __builtin_amdgcn_wave_barrier();
      build(dm * (attempt == 1 ? 1e-10 : 1e-6));
    }
    bad = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      if ((uint32_t)k >= R) continue; /* wave-uniform */
      if (h == (uint32_t)(k & 1)) rowk[j] = a[k >> 1];
      if (j == (uint32_t)k) {
#pragma unroll
        for (int q = 0; q < Q; ++q) colk[h + 2u * (uint32_t)q] = a[q];
      }
      __builtin_amdgcn_wave_barrier();
      double const p = rowk[k];
      bad |= !(p > 0.0);
      double const pinv = 1.0 / p;
      double const rk = ((j == (uint32_t)k) ? 1.0 : rowk[j]) * pinv;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        uint32_t const i = h + 2u * (uint32_t)q;
        double const f = colk[i];
        a[q] = (i == (uint32_t)k) ? rk : ((j == (uint32_t)k) ? 0.0 : a[q]) - f * rk;
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (!bad) break;
    reg = 1;
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    uint32_t const i = h + 2u * (uint32_t)q;
    if (i < R && j < R) Vinv[(size_t)i * R + j] = a[q];
  }
  if (threadIdx.x == 0 && status) {
    int const v = bad ? 2 : reg;
    if (v > *status) *status = v;
  }
}

__global__ __launch_bounds__(64) void k_solve_inv32(double const *__restrict__ G, uint32_t N,
                                                    uint32_t m, uint32_t R,
                                                    double *__restrict__ Vinv,
                                                    int *__restrict__ status)
{
  solve_inv32_body(G, N, m, R, Vinv, status);
}

template <int Q>
__global__ __launch_bounds__(DB) void k_solve_inv(double const *__restrict__ G,
                                                  uint32_t N, uint32_t m,
                                                  uint32_t R,
                                                  double *__restrict__ Vinv,
                                                  int *__restrict__ status)
{
  solve_inv_body<Q>(G, N, m, R, Vinv, status);
}

/* A = M * Vinv (FP64 accumulation), per-block column sum of squares / max */
template <typename T>
__device__ __forceinline__ void matmul_stats_body(uint32_t bid, uint32_t nblk, T const *__restrict__ M,
                                                     double const *__restrict__ Vinv,
                                                     T *__restrict__ A, uint64_t I,
                                                     uint32_t R,
                                                     double *__restrict__ part)
{
  HIP_DYNAMIC_SHARED(double, sv) /* R*R Vinv, then 2*DB reduction scratch */
  for (uint32_t x = threadIdx.x; x < R * R; x += DB) sv[x] = Vinv[x];
  double *red = sv + R * R;
  __syncthreads();
  uint32_t const rpi = DB / R; /* rows per iteration */
  uint32_t const rr = threadIdx.x / R, cc = threadIdx.x - rr * R;
  double ssq = 0.0, mx = -1.0e300;
  if (rr < rpi) {
    uint64_t const chunk = (I + nblk - 1) / nblk;
    uint64_t const lo = (uint64_t)bid * chunk;
    uint64_t const hi = lo + chunk < I ? lo + chunk : I;
    /* four rows at a time (independent product chains, each over j in
       order), their statistics taken in row order */
    uint64_t i = lo + rr;
    for (; i + 3u * rpi < hi; i += 4u * rpi) {
      T const *mrow = M + (size_t)i * R;
      size_t const st = (size_t)rpi * R;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (uint32_t j = 0; j < R; ++j) {
        double const vj = sv[j * R + cc];
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] += (double)mrow[u * st + j] * vj;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        T const o = (T)s[u];
        A[(size_t)(i + (uint64_t)u * rpi) * R + cc] = o;
        double const od = (double)o;
        ssq += od * od;
        mx = od > mx ? od : mx;
      }
    }
    for (; i < hi; i += rpi) {
      T const *mrow = M + (size_t)i * R;
      double s = 0.0;
      for (uint32_t j = 0; j < R; ++j) s += (double)mrow[j] * sv[j * R + cc];
      T const o = (T)s;
      A[(size_t)i * R + cc] = o;
      double const od = (double)o;
      ssq += od * od;
      mx = od > mx ? od : mx;
    }
  }
  red[threadIdx.x] = ssq;
  red[DB + threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x < R) {
    double s = 0.0, m = -1.0e300;
    for (uint32_t t = 0; t < rpi; ++t) {
      s += red[t * R + threadIdx.x];
      double const v = red[DB + t * R + threadIdx.x];
      m = v > m ? v : m;
    }
    part[(size_t)bid * 2 * R + threadIdx.x] = s;
    part[(size_t)bid * 2 * R + R + threadIdx.x] = m;
  }
}

template <typename T>
__global__ __launch_bounds__(DB) void k_matmul_stats(T const *__restrict__ M,
                                                     double const *__restrict__ Vinv,
                                                     T *__restrict__ A, uint64_t I,
                                                     uint32_t R,
                                                     double *__restrict__ part)
{
  matmul_stats_body<T>(blockIdx.x, gridDim.x, M, Vinv, A, I, R, part);
}

/* lambda from the column statistics: 2-norm (first sweep) or max(1, max).
 * Block: R columns x C = DB/R chunks of the partial rows (chunk c takes rows
 * c, c+C, ...), combined in chunk order (R <= 64). */
__device__ __forceinline__ void lambda_body(double const *__restrict__ part,
                                               uint32_t nblk, uint32_t R,
                                               uint32_t const *__restrict__ iter,
                                               double *__restrict__ lambda)
{
  __shared__ double rs[DB], rm[DB];
  uint32_t const C = DB / R;
  uint32_t const c = threadIdx.x / R, r = threadIdx.x - c * R;
  double s = 0.0, m = -1.0e300;
  if (c < C) {
    /* four partial rows' loads in flight, combined in row order */
    uint32_t b = c;
    for (; b + 3u * C < nblk; b += 4u * C) {
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        x[u] = part[(size_t)(b + u * C) * 2 * R + r];
        y[u] = part[(size_t)(b + u * C) * 2 * R + R + r];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s += x[u];
        m = y[u] > m ? y[u] : m;
      }
    }
    for (; b < nblk; b += C) {
      s += part[(size_t)b * 2 * R + r];
      double const v = part[(size_t)b * 2 * R + R + r];
      m = v > m ? v : m;
    }
  }
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x < R) {
    for (uint32_t q = 1; q < C; ++q) {
      s += rs[q * R + r];
      double const v = rm[q * R + r];
      m = v > m ? v : m;
    }
    int const first = (*iter == 0u);
    lambda[r] = first ? sqrt(s) : (m > 1.0 ? m : 1.0);
  }
}

__global__ __launch_bounds__(DB) void k_lambda(double const *__restrict__ part,
                                               uint32_t nblk, uint32_t R,
                                               uint32_t const *__restrict__ iter,
                                               double *__restrict__ lambda)
{
  lambda_body(part, nblk, R, iter, lambda);
}

/* per-block partial of sum_i A[i,r] * M[i,r] */
template <typename T>
__device__ __forceinline__ void inner_part_body(uint32_t bid, uint32_t nblk, T const *__restrict__ A,
                                                   T const *__restrict__ M,
                                                   uint64_t I, uint32_t R,
                                                   double *__restrict__ part)
{
  __shared__ double red[DB];
  uint32_t const rpi = DB / R;
  uint32_t const rr = threadIdx.x / R, cc = threadIdx.x - rr * R;
  double s = 0.0;
  if (rr < rpi) {
    uint64_t const chunk = (I + nblk - 1) / nblk;
    uint64_t const lo = (uint64_t)bid * chunk;
    uint64_t const hi = lo + chunk < I ? lo + chunk : I;
    uint64_t i = lo + rr;
    for (; i + 3u * rpi < hi; i += 4u * rpi) {
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        x[u] = (double)A[(size_t)(i + (uint64_t)u * rpi) * R + cc];
        y[u] = (double)M[(size_t)(i + (uint64_t)u * rpi) * R + cc];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) s += x[u] * y[u];
    }
    for (; i < hi; i += rpi)
      s += (double)A[(size_t)i * R + cc] * (double)M[(size_t)i * R + cc];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < R) {
    double t = 0.0;
    for (uint32_t q = 0; q < rpi; ++q) t += red[q * R + threadIdx.x];
    part[(size_t)bid * R + threadIdx.x] = t;
  }
}

template <typename T>
__global__ __launch_bounds__(DB) void k_inner_part(T const *__restrict__ A,
                                                   T const *__restrict__ M,
                                                   uint64_t I, uint32_t R,
                                                   double *__restrict__ part)
{
  inner_part_body<T>(blockIdx.x, gridDim.x, A, M, I, R, part);
}

/* per-block partial sum of squares */
template <typename T>
__global__ __launch_bounds__(DB) void k_sqnorm_part(T const *__restrict__ x,
                                                    uint64_t n,
                                                    double *__restrict__ part)
{
  __shared__ double red[DB];
  uint64_t const chunk = (n + gridDim.x - 1) / gridDim.x;
  uint64_t const lo = (uint64_t)blockIdx.x * chunk;
  uint64_t const hi = lo + chunk < n ? lo + chunk : n;
  double s = 0.0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += DB) {
    double const v = (double)x[i];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t k = DB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* fit = 1 - sqrt(|X|^2 + lambda^T (hada G) lambda - 2 <X,Z>) / |X|
 * (adatm_cpd.c:59-90, 109-160, 179-200); one block */
__device__ __forceinline__ void fit_final_body(double const *__restrict__ part,
                                                  uint32_t nblk,
                                                  double const *__restrict__ lambda,
                                                  double const *__restrict__ G,
                                                  uint32_t N, uint32_t R,
                                                  double ttnormsq,
                                                  double *__restrict__ fits,
                                                  uint32_t fcap,
                                                  uint32_t *__restrict__ iter)
{
  __shared__ double red[DB];
  __shared__ double red2[DB];
  double inner = 0.0;
  for (uint32_t r = threadIdx.x; r < R; r += DB) {
    double s = 0.0;
    uint32_t b = 0;
    for (; b + 4u <= nblk; b += 4u) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = part[(size_t)(b + u) * R + r];
#pragma unroll
      for (int u = 0; u < 4; ++u) s += x[u];
    }
    for (; b < nblk; ++b) s += part[(size_t)b * R + r];
    inner += s * lambda[r];
  }
  double nrm = 0.0;
  for (uint32_t x = threadIdx.x; x < R * R; x += DB) {
    uint32_t const i = x / R, j = x - i * R;
    double h = 1.0;
    for (uint32_t k = 0; k < N; ++k) h *= G[(size_t)k * R * R + x];
    nrm += h * lambda[i] * lambda[j];
  }
  red[threadIdx.x] = inner;
  red2[threadIdx.x] = nrm;
  __syncthreads();
  for (uint32_t s = DB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x] += red[threadIdx.x + s];
      red2[threadIdx.x] += red2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double const norm_mats = fabs(red2[0]);
    double const residual = sqrt(ttnormsq + norm_mats - 2.0 * red[0]);
    uint32_t const it = *iter;
    if (it < fcap) fits[it] = 1.0 - residual / sqrt(ttnormsq); /* never past the buffer */
    *iter = it + 1u;
  }
}

__global__ __launch_bounds__(DB) void k_fit_final(double const *__restrict__ part,
                                                  uint32_t nblk,
                                                  double const *__restrict__ lambda,
                                                  double const *__restrict__ G,
                                                  uint32_t N, uint32_t R,
                                                  double ttnormsq,
                                                  double *__restrict__ fits,
                                                  uint32_t fcap,
                                                  uint32_t *__restrict__ iter)
{
  fit_final_body(part, nblk, lambda, G, N, R, ttnormsq, fits, fcap, iter);
}

/* blocks of the row-parallel dense steps: at least `rows` rows each, at most
 * NBLK_MAX blocks (the partial-sum rows) */
uint32_t rows_blocks(uint64_t I, uint32_t rows)
{
  uint64_t const b = (I + rows - 1) / rows;
  return (uint32_t)(b < 1 ? 1 : (b > NBLK_MAX ? NBLK_MAX : b));
}

uint32_t dense_blocks(uint64_t I)
{
  uint64_t const b = (I + 255) / 256;
  return (uint32_t)(b < 1 ? 1 : (b > NBLK_MAX ? NBLK_MAX : b));
}

template <typename T>
hipError_t gram_impl(hipStream_t s, T *A, bf16raw *Abf, uint64_t I, uint32_t R,
                     double const *lambda, double *part, double *G_out)
{
  uint32_t const nb = rows_blocks(I, GRAM_ROWS);
  if constexpr (sizeof(T) == 4) {
    hipLaunchKernelGGL(k_gram_mfma, dim3(nb), dim3(DB), 0, s, (float *)A, Abf, I, R,
                       lambda, part);
  } else if (R <= 16) {
    hipLaunchKernelGGL((k_gram_mfma64<1>), dim3(nb), dim3(DB), 0, s, A, I, R, lambda, part);
  } else if (R <= 32) {
    hipLaunchKernelGGL((k_gram_mfma64<2>), dim3(nb), dim3(DB), 0, s, A, I, R, lambda, part);
  } else if (R <= 48) {
    hipLaunchKernelGGL((k_gram_mfma64<3>), dim3(nb), dim3(DB), 0, s, A, I, R, lambda, part);
  } else {
    hipLaunchKernelGGL((k_gram_mfma64<4>), dim3(nb), dim3(DB), 0, s, A, I, R, lambda, part);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  uint32_t const len = R * R;
  hipLaunchKernelGGL(k_sum_parts, dim3((len + 31) / 32), dim3(DB), 0, s,
                     part, nb, len, G_out);
  return hipGetLastError();
}

/* ---- grouped dense updates (one descriptor per instance) ----------------- */
template <int Q>
__global__ __launch_bounds__(DB) void g_solve(adatm_kdense const *__restrict__ d)
{
  adatm_kdense const &x = d[blockIdx.x];
  solve_inv_body<Q>(x.G, x.N, x.m, x.R, x.Vinv, x.status);
}

__global__ __launch_bounds__(64) void g_solve32(adatm_kdense const *__restrict__ d)
{
  adatm_kdense const &x = d[blockIdx.x];
  solve_inv32_body(x.G, x.N, x.m, x.R, x.Vinv, x.status);
}

template <typename T>
__global__ __launch_bounds__(DB) void g_matmul(adatm_kdense const *__restrict__ d,
                                               uint32_t const *__restrict__ bm)
{
  adatm_kdense const &x = d[bm[blockIdx.x]];
  matmul_stats_body<T>(blockIdx.x - x.mm0, x.nmm, (T const *)x.M, x.Vinv, (T *)x.A, x.I,
                       x.R, x.part);
}

__global__ __launch_bounds__(DB) void g_lambda(adatm_kdense const *__restrict__ d)
{
  adatm_kdense const &x = d[blockIdx.x];
  lambda_body(x.part, x.nmm, x.R, x.iter, x.lambda);
}

__global__ __launch_bounds__(DB) void g_gram_mfma(adatm_kdense const *__restrict__ d,
                                                  uint32_t const *__restrict__ bm)
{
  adatm_kdense const &x = d[bm[blockIdx.x]];
  gram_mfma_body(blockIdx.x - x.gr0, x.ngr, (float *)x.A, (bf16raw *)x.Abf, x.I, x.R,
                 x.lambda, x.part);
}

template <int NT>
__global__ __launch_bounds__(DB) void g_gram_mfma64(adatm_kdense const *__restrict__ d,
                                                    uint32_t const *__restrict__ bm)
{
  adatm_kdense const &x = d[bm[blockIdx.x]];
  gram_mfma64_body<NT>(blockIdx.x - x.gr0, x.ngr, (double *)x.A, x.I, x.R, x.lambda,
                       x.part);
}

__global__ __launch_bounds__(DB) void g_sum(adatm_kdense const *__restrict__ d,
                                            uint32_t const *__restrict__ bm)
{
  adatm_kdense const &x = d[bm[blockIdx.x]];
  sum_parts_body(blockIdx.x - x.sp0, x.part, x.ngr, x.R * x.R, x.G_out);
}

template <typename T>
__global__ __launch_bounds__(DB) void g_inner(adatm_kdense const *__restrict__ d,
                                              uint32_t const *__restrict__ bm)
{
  adatm_kdense const &x = d[bm[blockIdx.x]];
  inner_part_body<T>(blockIdx.x - x.in0, x.nin, (T const *)x.A, (T const *)x.M, x.I, x.R,
                     x.part);
}

__global__ __launch_bounds__(DB) void g_fit(adatm_kdense const *__restrict__ d)
{
  adatm_kdense const &x = d[blockIdx.x];
  fit_final_body(x.part, x.nin, x.lambda, x.G, x.N, x.R, x.ttnormsq, x.fits, x.fcap, x.iter);
}

} // namespace

/* ======================================================================== */
extern "C" {

static int herr(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

int adatm_k_zero(hipStream_t s, void *p, size_t nbytes)
{
  if (nbytes == 0) return 0;
  if ((((uintptr_t)p) & 3u) == 0 && (nbytes & 3u) == 0) {
    size_t const nw = nbytes / 4;
    size_t b = (nw + 255) / 256;
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_zero_words, dim3((uint32_t)b), dim3(256), 0, s,
                       (uint32_t *)p, nw);
  } else {
    hipLaunchKernelGGL(k_zero_bytes, dim3((uint32_t)((nbytes + 255) / 256)),
                       dim3(256), 0, s, (uint8_t *)p, nbytes);
  }
  return herr(hipGetLastError());
}

size_t adatm_k_part_doubles(uint32_t R)
{
  size_t const a = (size_t)R * R, b = 2 * (size_t)R;
  return (size_t)NBLK_MAX * (a > b ? a : b) + 64;
}

int adatm_k_gram(hipStream_t s, int prec, void const *A, uint64_t I, uint32_t R,
                 double *part, double *G_out)
{
  hipError_t e;
  if (prec == KP_FP64)
    e = gram_impl<double>(s, (double *)A, nullptr, I, R, nullptr, part, G_out);
  else
    e = gram_impl<float>(s, (float *)A, nullptr, I, R, nullptr, part, G_out);
  return herr(e);
}

int adatm_k_sqnorm(hipStream_t s, int prec, void const *x, uint64_t n,
                   double *part, double *out)
{
  uint32_t const nb = NBLK_MAX;
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_sqnorm_part<double>), dim3(nb), dim3(DB), 0, s,
                       (double const *)x, n, part);
  else
    hipLaunchKernelGGL((k_sqnorm_part<float>), dim3(nb), dim3(DB), 0, s,
                       (float const *)x, n, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(DB), 0, s, part, nb, 1u, out);
  return herr(hipGetLastError());
}

int adatm_k_solve_inv(hipStream_t s, double const *G, uint32_t N, uint32_t m,
                      uint32_t R, double *Vinv, int *status)
{
  if (R > 64) return herr(hipErrorInvalidValue);
  if (R <= 32)
    hipLaunchKernelGGL(k_solve_inv32, dim3(1), dim3(64), 0, s, G, N, m, R, Vinv, status);
  else
    hipLaunchKernelGGL((k_solve_inv<16>), dim3(1), dim3(DB), 0, s, G, N, m, R, Vinv, status);
  return herr(hipGetLastError());
}

void adatm_k_dense_blocks(uint64_t I, uint32_t R, uint32_t *nmm, uint32_t *ngr,
                          uint32_t *nsp)
{
  *nmm = rows_blocks(I, DB / R);
  *ngr = rows_blocks(I, GRAM_ROWS);
  *nsp = (R * R + 31) / 32;
}

int adatm_k_dense_grouped(hipStream_t s, int prec, adatm_kdense const *desc, uint32_t n,
                          uint32_t Rmax, uint32_t const *bm_mm, uint32_t nb_mm,
                          uint32_t const *bm_gr, uint32_t nb_gr, uint32_t const *bm_sp,
                          uint32_t nb_sp)
{
  if (n == 0) return 0;
  if (Rmax > 64) return herr(hipErrorInvalidValue);
  if (Rmax <= 32)
    hipLaunchKernelGGL(g_solve32, dim3(n), dim3(64), 0, s, desc);
  else
    hipLaunchKernelGGL((g_solve<16>), dim3(n), dim3(DB), 0, s, desc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  size_t const lds = ((size_t)Rmax * Rmax + 2 * DB) * sizeof(double);
  if (prec == KP_FP64)
    hipLaunchKernelGGL((g_matmul<double>), dim3(nb_mm), dim3(DB), lds, s, desc, bm_mm);
  else
    hipLaunchKernelGGL((g_matmul<float>), dim3(nb_mm), dim3(DB), lds, s, desc, bm_mm);
  if ((e = hipGetLastError()) != hipSuccess) return herr(e);
  hipLaunchKernelGGL(g_lambda, dim3(n), dim3(DB), 0, s, desc);
  if ((e = hipGetLastError()) != hipSuccess) return herr(e);
  if (prec == KP_FP64) {
    if (Rmax <= 16)
      hipLaunchKernelGGL((g_gram_mfma64<1>), dim3(nb_gr), dim3(DB), 0, s, desc, bm_gr);
    else if (Rmax <= 32)
      hipLaunchKernelGGL((g_gram_mfma64<2>), dim3(nb_gr), dim3(DB), 0, s, desc, bm_gr);
    else if (Rmax <= 48)
      hipLaunchKernelGGL((g_gram_mfma64<3>), dim3(nb_gr), dim3(DB), 0, s, desc, bm_gr);
    else
      hipLaunchKernelGGL((g_gram_mfma64<4>), dim3(nb_gr), dim3(DB), 0, s, desc, bm_gr);
  } else {
    hipLaunchKernelGGL(g_gram_mfma, dim3(nb_gr), dim3(DB), 0, s, desc, bm_gr);
  }
  if ((e = hipGetLastError()) != hipSuccess) return herr(e);
  hipLaunchKernelGGL(g_sum, dim3(nb_sp), dim3(DB), 0, s, desc, bm_sp);
  return herr(hipGetLastError());
}

uint32_t adatm_k_fit_blocks(uint64_t I) { return dense_blocks(I); }

int adatm_k_fit_grouped(hipStream_t s, int prec, adatm_kdense const *desc, uint32_t n,
                        uint32_t const *bm_in, uint32_t nb_in)
{
  if (n == 0) return 0;
  if (prec == KP_FP64)
    hipLaunchKernelGGL((g_inner<double>), dim3(nb_in), dim3(DB), 0, s, desc, bm_in);
  else
    hipLaunchKernelGGL((g_inner<float>), dim3(nb_in), dim3(DB), 0, s, desc, bm_in);
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  hipLaunchKernelGGL(g_fit, dim3(n), dim3(DB), 0, s, desc);
  return herr(hipGetLastError());
}

int adatm_k_update(hipStream_t s, int prec, void const *M, double const *Vinv,
                   void *A, void *Abf, uint64_t I, uint32_t R,
                   uint32_t const *iter, double *lambda, double *part,
                   double *G_out)
{
  uint32_t const nb = rows_blocks(I, DB / R);
  size_t const lds = ((size_t)R * R + 2 * DB) * sizeof(double);
  hipError_t e;
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_matmul_stats<double>), dim3(nb), dim3(DB), lds, s,
                       (double const *)M, Vinv, (double *)A, I, R, part);
  else
    hipLaunchKernelGGL((k_matmul_stats<float>), dim3(nb), dim3(DB), lds, s,
                       (float const *)M, Vinv, (float *)A, I, R, part);
  e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  hipLaunchKernelGGL(k_lambda, dim3(1), dim3(DB), 0, s, part, nb, R, iter, lambda);
  e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  if (prec == KP_FP64)
    e = gram_impl<double>(s, (double *)A, nullptr, I, R, lambda, part, G_out);
  else
    e = gram_impl<float>(s, (float *)A, (bf16raw *)Abf, I, R, lambda, part,
                         G_out);
  return herr(e);
}

int adatm_k_fit(hipStream_t s, int prec, void const *A_last, void const *M,
                uint64_t I_last, double const *lambda, double const *G,
                uint32_t N, uint32_t R, double ttnormsq, double *part,
                double *fits, uint32_t fcap, uint32_t *iter)
{
  uint32_t const nb = dense_blocks(I_last);
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_inner_part<double>), dim3(nb), dim3(DB), 0, s,
                       (double const *)A_last, (double const *)M, I_last, R,
                       part);
  else
    hipLaunchKernelGGL((k_inner_part<float>), dim3(nb), dim3(DB), 0, s,
                       (float const *)A_last, (float const *)M, I_last, R,
                       part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return herr(e);
  hipLaunchKernelGGL(k_fit_final, dim3(1), dim3(DB), 0, s, part, nb, lambda, G,
                     N, R, ttnormsq, fits, fcap, iter);
  return herr(hipGetLastError());
}

/* the FP32 -> BF16 rounding of the BF16 factor-row copies, on the host */
uint16_t adatm_k_bf16_bits(float f) { return to_bf16(f); }

int adatm_k_convert_in(hipStream_t s, int prec, double const *src64, void *dst,
                       void *dst_bf, uint64_t n)
{
  uint64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_convert_in<double>), dim3((uint32_t)b), dim3(256), 0,
                       s, src64, (double *)dst, (bf16raw *)nullptr, n);
  else
    hipLaunchKernelGGL((k_convert_in<float>), dim3((uint32_t)b), dim3(256), 0,
                       s, src64, (float *)dst, (bf16raw *)dst_bf, n);
  return herr(hipGetLastError());
}

int adatm_k_convert_out(hipStream_t s, int prec, void const *src, double *dst64,
                        uint64_t n)
{
  uint64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_convert_out<double>), dim3((uint32_t)b), dim3(256), 0,
                       s, (double const *)src, dst64, n);
  else
    hipLaunchKernelGGL((k_convert_out<float>), dim3((uint32_t)b), dim3(256), 0,
                       s, (float const *)src, dst64, n);
  return herr(hipGetLastError());
}

} /* extern "C" */
