/*
 * Internal declarations shared by the host C sources of libsplatt_adatm.
 */
#ifndef ADATM_HOST_H
#define ADATM_HOST_H

#include <stdint.h>
#include <stddef.h>

#include "../../../include/splatt.h"
#include "../../../include/adatm.h"

#define ADATM_NONE 0xFFFFFFFFu

/* thread-local last error message */
void adatm_set_error(char const *fmt, ...);

/* ---- host sorting / structure (coo.c) ---------------------------------- */

/* perm (nnz entries) := stable argsort of the coordinates, lexicographic by
 * modes order[0], order[1], ..., order[nkeys-1]. */
int adatm_lex_sort_perm(adatm_coo const *t, uint32_t const *order,
                        uint32_t nkeys, uint32_t *perm);

/* Group CSF in host memory (u32 structure).  Level l holds F[l] nodes;
 * fids[l][n] is the coordinate of mode order[l] of node n (fids[0][n] == n
 * after canonicalisation), fptr[l][n] the first child (l < N-1, F[l]+1
 * entries), par[l][n] the parent at level l-1 (l >= 1). */
typedef struct {
  uint32_t N;
  uint64_t nnz;
  uint32_t order[ADATM_MAX_NMODES];
  uint64_t F[ADATM_MAX_NMODES];
  uint32_t *fptr[ADATM_MAX_NMODES];
  uint32_t *fids[ADATM_MAX_NMODES];
  uint32_t *par[ADATM_MAX_NMODES];
  double *vals;
} adatm_hcsf;

int adatm_hcsf_build(adatm_coo const *t, uint32_t const *order,
                     adatm_hcsf *out);
void adatm_hcsf_free(adatm_hcsf *c);

/* ---- planner (planner.c) ----------------------------------------------- */
int adatm_plan_enumerate(uint32_t nmodes, uint64_t const *dims,
                         adatm_plan *cands, int maxc);

/* ---- engine internals shared with the SPLATT layer (engine.c) ---------- */
/* device-built engine with a given mode -> (group, tag) map */
int adatm_engine_create_use(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                            int precision, int device, uint32_t const *use_grp,
                            uint32_t const *use_tag, adatm_engine **out);
/* one CSF (sorted values, fptr, fids of every level) built on the device */
int adatm_hcsf_build_device(adatm_coo const *t, uint32_t const *order, int device,
                            adatm_hcsf *h);
uint32_t adatm_engine_nmodes(adatm_engine const *e);
uint64_t adatm_engine_dim(adatm_engine const *e, uint32_t m);
uint32_t adatm_engine_rank(adatm_engine const *e);
int adatm_engine_precision(adatm_engine const *e);
int adatm_count_fibers_host(adatm_coo const *t, uint32_t const *order,
                            uint64_t *nfibs);
int adatm_coo_from_sptensor(sptensor_t const *tt, adatm_coo *t);
/* analytic model of one plan from every group's fiber counts (no quirks) */
void adatm_plan_model(uint32_t N, uint64_t nnz, uint64_t R, adatm_plan const *P,
                      uint64_t const *Fg, uint64_t *pspace, uint64_t *pops);
/* fiber counts of norders level orders (F: norders x N) in one device pass */
int adatm_count_fibers_device_many(adatm_coo const *t, uint32_t const *orders,
                                   uint32_t norders, uint64_t *F, int device);

/* sum of squares in a fixed association (thread-count independent) */
double adatm_sumsq(double const *v, uint64_t n);

/* ---- seconds since an arbitrary origin --------------------------------- */
double adatm_wtime(void);

#endif
