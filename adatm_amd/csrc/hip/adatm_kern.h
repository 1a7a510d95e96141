/*
 * C ABI of the HIP kernels of libsplatt_adatm (launchers, one per kernel).
 * Internal to the library: the host C engine calls these; users call the
 * SPLATT / adatm_* entry points of include/.
 */
#ifndef ADATM_KERN_H
#define ADATM_KERN_H

#include <stdint.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* value layouts of the device path */
enum { KP_FP64 = 0, KP_FP32 = 1, KP_BF16ROWS = 2 };

#define ADATM_KMAXL 16

/* segmented-reduction work list (see adatm_segwork in the host code) */
typedef struct {
  uint32_t nitems;
  uint32_t nheavy;
  uint32_t const *start; /* nitems+1 */
  uint32_t const *seg;   /* nitems+1 */
  uint32_t const *slot;  /* nitems */
  uint32_t const *hseg;  /* nheavy */
  uint32_t const *hslot; /* nheavy+1 */
} adatm_kwork;

/* out[p] = sum_{c in [ptr[p],ptr[p+1])} in(c) o U[ids[c]]; in(c) is the
 * scalar xin[c] when leaf, else the row xin[c*R ...].  Heavy pieces write
 * carry slots, then the fix-up sums them. */
int adatm_k_seg_reduce(hipStream_t s, int prec, int leaf, adatm_kwork const *w,
                       uint32_t const *ptr, uint32_t const *ids,
                       void const *xin, void const *U, void *out, void *carry,
                       uint32_t R);

/* Reuse MTTKRP at CSF level d (d >= 1) over level-d nodes sorted by id:
 * out[i] = sum_{q in [rowptr[i],rowptr[i+1])} Y(par_d[n]) o Z[n], n=perm[q],
 * Y(p) = prod_{l<d} U_l[id_l(ancestor_l)], id_0 = node ordinal. */
typedef struct {
  uint32_t d;
  uint32_t const *par[ADATM_KMAXL]; /* par[l]: level l -> l-1, l = 1..d */
  uint32_t const *ids[ADATM_KMAXL]; /* ids[l], l = 1..d-1 */
  void const *U[ADATM_KMAXL];       /* U[l] = factor of the level-l mode */
} adatm_kpath;

int adatm_k_scatter(hipStream_t s, int prec, adatm_kwork const *w,
                    uint32_t const *rowptr, uint32_t const *perm,
                    adatm_kpath const *path,
                    void const *Z, void *out, void *carry, uint32_t R);

/* zero n bytes with a kernel (graph-capture safe) */
int adatm_k_zero(hipStream_t s, void *p, size_t nbytes);

/* Dense CP-ALS steps (FP64 arithmetic; factors/M in the path's precision).
 * G: N grams of R x R (FP64) contiguous.  part: scratch of
 * adatm_k_part_doubles(R) doubles. */
size_t adatm_k_part_doubles(uint32_t R);
int adatm_k_gram(hipStream_t s, int prec, void const *A, uint64_t I,
                 uint32_t R, double *part, double *G_out);
/* sum of squares of n values (FP64 accumulation) into *out */
int adatm_k_sqnorm(hipStream_t s, int prec, void const *x, uint64_t n,
                   double *part, double *out);
int adatm_k_solve_inv(hipStream_t s, double const *G, uint32_t N, uint32_t m,
                      uint32_t R, double *Vinv, int *status);
/* A = M Vinv, then lambda (2-norm when *iter == 0, else max(1, colmax)),
 * A /= lambda, G_out = A^T A, bf16 copy refreshed when Abf != NULL.
 * `iter` is a device counter so that one captured graph serves all sweeps. */
int adatm_k_update(hipStream_t s, int prec, void const *M, double const *Vinv,
                   void *A, void *Abf, uint64_t I, uint32_t R,
                   uint32_t const *iter, double *lambda, double *part,
                   double *G_out);
/* fit of the current model into fits[*iter], then ++*iter (needs M of the
 * last mode). */
int adatm_k_fit(hipStream_t s, int prec, void const *A_last, void const *M,
                uint64_t I_last, double const *lambda, double const *G,
                uint32_t N, uint32_t R, double ttnormsq, double *part,
                double *fits, uint32_t *iter);
/* FP64 host matrix <-> device factor in the path's precision */
int adatm_k_convert_in(hipStream_t s, int prec, double const *src64,
                       void *dst, void *dst_bf, uint64_t n);
int adatm_k_convert_out(hipStream_t s, int prec, void const *src,
                        double *dst64, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif
