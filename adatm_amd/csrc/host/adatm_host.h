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

/* Work decomposition of a segmented reduction: children [0, nchild) in
 * segments seg p = [ptr[p], ptr[p+1]).  Light items cover whole segments
 * seg[k] .. seg[k+1]-1 with at most `tile` children; a segment longer than
 * `tile` is cut into pieces of `tile` children, each writing a carry slot
 * that a fix-up pass sums in piece order (deterministic). */
typedef struct {
  uint32_t nitems;
  uint32_t nheavy;
  uint32_t nslots;
  uint32_t *start; /* nitems+1: first child of item */
  uint32_t *seg;   /* nitems+1: first segment of item */
  uint32_t *slot;  /* nitems: carry slot of a heavy piece, ADATM_NONE else */
  uint32_t *hseg;  /* nheavy: the split segment */
  uint32_t *hslot; /* nheavy+1: its slots [hslot[h], hslot[h+1]) */
} adatm_segwork;

int adatm_segwork_build(uint32_t const *ptr, uint64_t nseg, uint32_t tile,
                        adatm_segwork *w);
void adatm_segwork_free(adatm_segwork *w);

/* Counting sort of node ids (values < nrows): perm = stable order by id,
 * rowptr (nrows+1) = segment of each id in perm. */
int adatm_sort_by_id(uint32_t const *ids, uint64_t n, uint64_t nrows,
                     uint32_t *perm, uint32_t *rowptr);

/* ---- planner (planner.c) ----------------------------------------------- */
int adatm_plan_enumerate(uint32_t nmodes, uint64_t const *dims,
                         adatm_plan *cands, int maxc);

/* ---- engine internals shared with the SPLATT layer (engine.c) ---------- */
int adatm_engine_create_hcsf(adatm_hcsf const *groups, uint32_t N,
                             uint64_t const *dims, uint64_t nnz,
                             double ttnormsq, uint64_t rank,
                             adatm_plan const *plan, int precision, int device,
                             uint32_t const *use_grp, uint32_t const *use_tag,
                             adatm_engine **out);
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
