/*
 * Memoized-MTTKRP kernels for MI355X (gfx950, wave64).
 *
 * A lane group of G lanes owns one work item; lane j holds V consecutive
 * columns of every R-wide row (G = R/V: V = 1, or in the wide layout 16
 * bytes per lane -- 4 FP32/BF16 or 2 FP64 columns), so each factor-row
 * gather and each memo-row access of a group is one contiguous R*sizeof(T)
 * transaction, made of 128-bit lane accesses in the wide layout.  R <= 64;
 * 64/G groups per wavefront.  Work items are balanced by the number of children they walk;
 * a parent longer than one item is cut into pieces whose partial rows go to
 * carry slots, summed in a fixed order by k_fix_chunk / k_fix_multi, so results are
 * deterministic and no output needs zero-filling or atomics.
 *
 *   k_reduce<K>  Z_top(p) = sum over the level-`top+K` descendants c of p of
 *                in(c) o U_K[id(c)], folded up level by level with the
 *                factor rows of the intermediate nodes.  K = 1 is the leaf
 *                SpTTM (p_spttm, src/mttkrp.c:891-928) or one Hadamard-reduce
 *                (p_spt2t_add_hada / p_spt2m_add_hada, src/mttkrp.c:828-888);
 *                K > 1 fuses K such passes so that the levels in between --
 *                the reference's temporaries tmp_vals / tmp_vals_2,
 *                src/mttkrp.c:991-1066 -- never reach HBM.  Per-level running
 *                sums live in registers, as p_propagate_up keeps them in
 *                per-level buffers (src/mttkrp.c:241-304).
 *   k_scatter<D> reuse MTTKRP of CSF level D (p_csf_mttkrp_internal_reuse_
 *                adaptive, src/mttkrp.c:1547-1682): out[i] = sum over level-D
 *                nodes n with id i of Y(n) o Z_D(n), Y(n) = product of the
 *                factor rows of n's ancestors.  The reference scatters under
 *                1024 striped locks; here nodes are pre-sorted by output row
 *                with their ancestors' row ids, so every output row is one
 *                segment of a segmented reduction.
 *   k_fix_chunk  merge of split parents: a lane group per chunk of <= 32 carry
 *                rows; k_fix_multi sums the chunk rows of parents with several.
 *
 * This translation unit holds the C entry points; the kernels of each value
 * type are compiled in mttkrp_fp64.hip / mttkrp_fp32.hip / mttkrp_bf16.hip
 * from the templates of mttkrp_impl.hpp, and the split-parent merge here.
 */
#include "mttkrp_impl.hpp"

/* the per-type launchers (mttkrp_fp64.hip, mttkrp_fp32.hip, mttkrp_bf16.hip) */
#define ADATM_MTTKRP_DECL(NAME)                                                     \
  hipError_t adatm_reduce_##NAME(hipStream_t s, int leaf, adatm_kwork const *w,    \
                                 adatm_kchain const *ch, void const *xin, void *out, \
                                 void *carry, void *carry2, uint32_t R);           \
  hipError_t adatm_scatter_##NAME(hipStream_t s, adatm_kwork const *w,             \
                                  uint32_t const *rowptr, adatm_kscat const *sc,  \
                                  void const *Z, int zs, void *out, void *carry,  \
                                  void *carry2, uint32_t R);
extern "C" {
ADATM_MTTKRP_DECL(fp64)
ADATM_MTTKRP_DECL(fp32)
ADATM_MTTKRP_DECL(bf16)
}

extern "C" {

int adatm_k_reduce(hipStream_t s, int prec, int leaf, adatm_kwork const *w,
                   adatm_kchain const *ch, void const *xin, void *out,
                   void *carry, void *carry2, uint32_t R)
{
  if (ch->K < 1 || ch->K > ADATM_KMAXFUSE || R < 1 || R > 64)
    return (int)hipErrorInvalidValue;
  hipError_t const e =
      (prec == KP_FP64)   ? adatm_reduce_fp64(s, leaf, w, ch, xin, out, carry, carry2, R)
      : (prec == KP_FP32) ? adatm_reduce_fp32(s, leaf, w, ch, xin, out, carry, carry2, R)
                          : adatm_reduce_bf16(s, leaf, w, ch, xin, out, carry, carry2, R);
  return e == hipSuccess ? 0 : (int)e;
}

int adatm_k_fixup(hipStream_t s, int prec, adatm_kwork const *w,
                  void const *carry, void *carry2, void *out, uint32_t R)
{
  if (R < 1 || R > 64) return (int)hipErrorInvalidValue;
  hipError_t const e = (prec == KP_FP64)
                           ? launch_fixup<double>(s, w, carry, carry2, out, R)
                           : launch_fixup<float>(s, w, carry, carry2, out, R);
  return e == hipSuccess ? 0 : (int)e;
}

int adatm_k_scatter(hipStream_t s, int prec, adatm_kwork const *w,
                    uint32_t const *rowptr, adatm_kscat const *sc,
                    void const *Z, int zscalar, void *out, void *carry,
                    void *carry2, uint32_t R)
{
  if (sc->d < 1 || sc->d > ADATM_KMAXSCAT || R < 1 || R > 64)
    return (int)hipErrorInvalidValue;
  hipError_t const e =
      (prec == KP_FP64)
          ? adatm_scatter_fp64(s, w, rowptr, sc, Z, zscalar, out, carry, carry2, R)
      : (prec == KP_FP32)
          ? adatm_scatter_fp32(s, w, rowptr, sc, Z, zscalar, out, carry, carry2, R)
          : adatm_scatter_bf16(s, w, rowptr, sc, Z, zscalar, out, carry, carry2, R);
  return e == hipSuccess ? 0 : (int)e;
}

} /* extern "C" */

/* ---- persistent launch ------------------------------------------------------ */
#define ADATM_PERSIST_DECL(NAME)                                                    \
  hipError_t adatm_persist_##NAME(hipStream_t s, uint32_t grid, adatm_kstage const *st, \
                                  uint64_t const *tasks, uint32_t const *qoff,     \
                                  uint32_t const *xmap, uint32_t nq, uint32_t *ctr, \
                                  uint32_t *stat, uint32_t spin);                  \
  hipError_t adatm_persist_grid_##NAME(int device, uint32_t *grid);                   \
  hipError_t adatm_grouped_##NAME(hipStream_t s, uint32_t type, uint32_t leaf,        \
                                  uint32_t K, uint32_t side, uint32_t d, uint32_t zs, \
                                  adatm_kstage const *st, uint64_t const *bmap,       \
                                  uint32_t nb);
extern "C" {
ADATM_PERSIST_DECL(fp64)
ADATM_PERSIST_DECL(fp32)
ADATM_PERSIST_DECL(bf16)
}

namespace {
__global__ void k_xcc_probe(uint32_t *out)
{
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}
} // namespace

extern "C" {

int adatm_k_persist_grid(int prec, int device, uint32_t *grid)
{
  hipError_t const e = (prec == KP_FP64)   ? adatm_persist_grid_fp64(device, grid)
                       : (prec == KP_FP32) ? adatm_persist_grid_fp32(device, grid)
                                           : adatm_persist_grid_bf16(device, grid);
  return e == hipSuccess ? 0 : (int)e;
}

int adatm_k_persist(hipStream_t s, int prec, uint32_t grid, adatm_kstage const *stages,
                    uint64_t const *tasks, uint32_t const *qoff, uint32_t const *xmap,
                    uint32_t nq, uint32_t *ctr, uint32_t *stat, uint32_t spin)
{
  if (grid == 0 || nq == 0 || nq > 16 || spin == 0) return (int)hipErrorInvalidValue;
  hipError_t const e =
      (prec == KP_FP64)
          ? adatm_persist_fp64(s, grid, stages, tasks, qoff, xmap, nq, ctr, stat, spin)
      : (prec == KP_FP32)
          ? adatm_persist_fp32(s, grid, stages, tasks, qoff, xmap, nq, ctr, stat, spin)
          : adatm_persist_bf16(s, grid, stages, tasks, qoff, xmap, nq, ctr, stat, spin);
  return e == hipSuccess ? 0 : (int)e;
}

int adatm_k_grouped(hipStream_t s, int prec, adatm_kstage const *kind,
                    adatm_kstage const *stages, uint64_t const *bmap, uint32_t nblocks)
{
  if (!kind || (kind->type == KS_REDUCE && (kind->K < 1 || kind->K > 5)) ||
      (kind->type == KS_SCATTER && (kind->d < 1 || kind->d > 5)))
    return (int)hipErrorInvalidValue;
  uint32_t const t = kind->type, l = kind->leaf, K = kind->K, sd = kind->side,
                 d = kind->d, z = kind->zs;
  hipError_t const e =
      (prec == KP_FP64)   ? adatm_grouped_fp64(s, t, l, K, sd, d, z, stages, bmap, nblocks)
      : (prec == KP_FP32) ? adatm_grouped_fp32(s, t, l, K, sd, d, z, stages, bmap, nblocks)
                          : adatm_grouped_bf16(s, t, l, K, sd, d, z, stages, bmap, nblocks);
  return e == hipSuccess ? 0 : (int)e;
}

int adatm_k_xcc_probe(hipStream_t s, uint32_t *out, uint32_t nblocks)
{
  if (nblocks == 0) return 0;
  hipLaunchKernelGGL(k_xcc_probe, dim3(nblocks), dim3(64), 0, s, out);
  hipError_t const e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

} /* extern "C" */
