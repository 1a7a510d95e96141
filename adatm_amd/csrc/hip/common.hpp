/*
 * Device helpers shared by the HIP translation units (gfx950, wave64).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NONE32 0xFFFFFFFFu

namespace {

/* bfloat16 factor-row storage: raw upper half of an FP32 */
struct bf16raw {
  uint16_t v;
};

__device__ __forceinline__ float ldu(float const *p) { return *p; }
__device__ __forceinline__ double ldu(double const *p) { return *p; }
__device__ __forceinline__ float ldu(bf16raw const *p)
{
  return __uint_as_float(((uint32_t)p->v) << 16);
}

/* value of v held by wave lane src (ds_bpermute: the source lane must be
 * active in the calling branch) */
__device__ __forceinline__ uint32_t lane_get(uint32_t v, uint32_t src)
{
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ float lane_get(float v, uint32_t src)
{
  return __uint_as_float(lane_get(__float_as_uint(v), src));
}
__device__ __forceinline__ double lane_get(double v, uint32_t src)
{
  uint32_t w[2];
  __builtin_memcpy(w, &v, 8);
  w[0] = lane_get(w[0], src);
  w[1] = lane_get(w[1], src);
  double r;
  __builtin_memcpy(&r, w, 8);
  return r;
}

/* value of v held by lane l of the wave, l wave-uniform (v_readlane: the
 * result is a scalar) */
__device__ __forceinline__ uint32_t lane_read(uint32_t v, uint32_t l)
{
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float lane_read(float v, uint32_t l)
{
  return __uint_as_float(lane_read(__float_as_uint(v), l));
}
__device__ __forceinline__ double lane_read(double v, uint32_t l)
{
  uint32_t w[2];
  __builtin_memcpy(w, &v, 8);
  w[0] = lane_read(w[0], l);
  w[1] = lane_read(w[1], l);
  double r;
  __builtin_memcpy(&r, w, 8);
  return r;
}

/* V consecutive row elements at p, loaded as one access of V*sizeof(TU)
 * bytes (p aligned to that size), converted to the compute type */
template <typename T, int V>
__device__ __forceinline__ void ldrow(T const *p, T (&o)[V])
{
  if constexpr (V == 1)
    o[0] = *p;
  else
    __builtin_memcpy(o, __builtin_assume_aligned(p, V * sizeof(T)), V * sizeof(T));
}
template <int V>
__device__ __forceinline__ void ldrow(bf16raw const *p, float (&o)[V])
{
  uint16_t r[V];
  if constexpr (V == 1)
    r[0] = p->v;
  else
    __builtin_memcpy(r, __builtin_assume_aligned(p, V * 2), V * 2);
#pragma unroll
  for (int v = 0; v < V; ++v) o[v] = __uint_as_float(((uint32_t)r[v]) << 16);
}
template <typename T, int V>
__device__ __forceinline__ void strow(T *p, T const (&o)[V])
{
  if constexpr (V == 1)
    *p = o[0];
  else
    __builtin_memcpy(__builtin_assume_aligned(p, V * sizeof(T)), o, V * sizeof(T));
}

/* Raw buffer access to a factor of less than 2^31 bytes (the descriptor's
 * record count; gfx9 buffer format word): a load at an offset at or past
 * the count returns zeros without touching memory. */
#define BUF_RECORDS 0x7fffffffu
#define BUF_NONE 0x80000000u /* an offset that loads nothing */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(void const *p)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)BUF_RECORDS,
                                           0x00020000);
}

/* ldrow from a buffer at byte offset off (V * sizeof(TU) bytes, one load) */
template <typename TU, typename T, int V>
__device__ __forceinline__ void ldrow_buf(__amdgpu_buffer_rsrc_t rs, uint32_t off, T (&o)[V])
{
  constexpr int B = V * (int)sizeof(TU);
  static_assert(B == 2 || B == 4 || B == 8 || B == 16, "row piece size");
  TU r[V];
  if constexpr (B == 16) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    u4 const t = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
    __builtin_memcpy(r, &t, 16);
  } else if constexpr (B == 8) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    u2 const t = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, 0);
    __builtin_memcpy(r, &t, 8);
  } else if constexpr (B == 4) {
    uint32_t const t = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
    __builtin_memcpy(r, &t, 4);
  } else {
    uint16_t const t = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)off, 0, 0);
    __builtin_memcpy(r, &t, 2);
  }
#pragma unroll
  for (int v = 0; v < V; ++v) o[v] = ldu(r + v);
}

/* strow with the nontemporal hint (global_store_dwordx4 ... nt for a 16-byte
 * piece): for rows written once and not read again by the writing pass, so
 * that they do not push the factor rows the pass gathers out of L2 */
template <typename T, int V>
__device__ __forceinline__ void strow_nt(T *p, T const (&o)[V])
{
  if constexpr (V * sizeof(T) == 16) {
    typedef T vec_t __attribute__((ext_vector_type(V)));
    vec_t v;
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = o[i];
    __builtin_nontemporal_store(v, (vec_t *)__builtin_assume_aligned(p, 16));
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) __builtin_nontemporal_store(o[i], p + i);
  }
}

/* FP32 -> BF16 with round-to-nearest-even.  Infinities keep their sign;
 * every NaN becomes a quiet NaN of its sign (its payload may sit in the low
 * 16 mantissa bits only, which truncation alone would turn into an
 * infinity). */
__host__ __device__ __forceinline__ uint16_t to_bf16(float f)
{
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

} // namespace
