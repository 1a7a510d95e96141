/*
 * Device engine: one memoization plan on one GPU.
 *
 * Holds, in HBM and for the life of the engine, every group's CSF (u32
 * structure, values in the path precision), the memoised semi-sparse
 * intermediates (vCSF values, adatm_rcsf.c:216-278), the factor matrices,
 * their Grams and the MTTKRP output.  Sequencing follows the reference's
 * dispatch (mttkrp.c:2037-2090):
 *   tag 0 (group root)   leaf SpTTM, then a Hadamard-reduce per level up to
 *                        the root (p_csf_mttkrp_root_genreuse_adaptive,
 *                        mttkrp.c:934-1079; p_csf_mttkrp_root,
 *                        mttkrp.c:766-824 when the group memoises nothing),
 *                        storing the levels b-1 .. b-n_imten;
 *   tag 1 (reuse)        from the memo level L = b - min(n_imten, b-d):
 *                        reduce L -> d if L > d, then the level-d scatter
 *                        (mttkrp.c:1547-1682).
 * Memo levels are overwritten by every root call (the reference adds into
 * them and never clears them, so its second and later sweeps are wrong for
 * N >= 4 -- adatm_rcsf.c:273-274 with mttkrp.c:967/1022/1035).
 * CP-ALS follows cpd_als_iterate_adaptive (adatm_cpd.c:256-374).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "adatm_host.h"
#include "../hip/adatm_kern.h"

#define MAXN ADATM_MAX_NMODES
#define RED_TILE 128u
#define SCAT_TILE 64u

typedef struct {
  adatm_kwork kw; /* device pointers */
  uint32_t nslots;
} dwork;

typedef struct {
  uint32_t memo_mode, b, nim;
  uint32_t order[MAXN];
  uint64_t F[MAXN];
  void *vals;
  uint32_t *ids[MAXN];
  uint32_t *ptr[MAXN];
  uint32_t *par[MAXN];
  dwork red[MAXN];
  void *memo[MAXN];
  uint32_t *sperm[MAXN];
  uint32_t *rowptr[MAXN];
  dwork scat[MAXN];
  uint32_t pmax; /* deepest level with a parent array */
} dgroup;

enum { KC_LEAF = 0, KC_LEVEL = 1, KC_SCATTER = 2, KC_DENSE = 3, KC_NCLASS = 4 };

struct adatm_engine {
  int device;
  hipStream_t stream;
  uint32_t N, R;
  int prec;
  size_t vsz; /* bytes per value of the path (memo, factor, M) */
  size_t wsz; /* bytes per gathered factor value */
  uint64_t nnz;
  uint64_t dims[MAXN];
  uint64_t maxdim;
  adatm_plan plan;
  uint32_t use_grp[MAXN], use_tag[MAXN], level_of[MAXN][MAXN];
  dgroup g[MAXN];
  void *A[MAXN];
  void *Abf[MAXN];
  void *M;
  double *G, *lambda, *Vinv, *part, *fits, *scal;
  uint32_t *iter;
  int *status;
  void *tmp[2];
  void *carry;
  size_t tmp_bytes, carry_bytes;
  double ttnormsq;
  size_t dev_bytes;
  uint32_t fits_cap;
  hipGraphExec_t gexec;
  hipGraph_t graph;
  int gvalid;
  /* optional per-class timing of launches */
  int instrument;
  hipEvent_t ev[2048];
  int ev_class[1024];
  double ev_bytes[1024];
  int nev;
};

/* ---- helpers ------------------------------------------------------------ */
static int hip_fail(hipError_t e, char const *what)
{
  adatm_set_error("%s: %s", what, hipGetErrorString(e));
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
             ? SPLATT_ERROR_NOMEMORY
             : SPLATT_ERROR_DEVICE;
}

#define HIPCHK(call, what)                                                     \
  do {                                                                         \
    hipError_t const e__ = (call);                                             \
    if (e__ != hipSuccess) return hip_fail(e__, what);                         \
  } while (0)

#define KCHK(call, what)                                                       \
  do {                                                                         \
    int const k__ = (call);                                                    \
    if (k__ != 0) return hip_fail((hipError_t)k__, what);                      \
  } while (0)

static int dalloc(adatm_engine *e, void **p, size_t bytes)
{
  *p = NULL;
  if (bytes == 0) return SPLATT_SUCCESS;
  hipError_t const r = hipMalloc(p, bytes);
  if (r != hipSuccess) {
    *p = NULL;
    adatm_set_error("hipMalloc(%zu bytes) failed: %s", bytes,
                    hipGetErrorString(r));
    return SPLATT_ERROR_NOMEMORY;
  }
  e->dev_bytes += bytes;
  return SPLATT_SUCCESS;
}

static int dupload(adatm_engine *e, void **p, void const *h, size_t bytes)
{
  int rc = dalloc(e, p, bytes);
  if (rc != SPLATT_SUCCESS) return rc;
  if (bytes) HIPCHK(hipMemcpy(*p, h, bytes, hipMemcpyHostToDevice), "upload");
  return SPLATT_SUCCESS;
}

static int upload_work(adatm_engine *e, adatm_segwork const *w, dwork *d)
{
  int rc;
  memset(d, 0, sizeof(*d));
  d->kw.nitems = w->nitems;
  d->kw.nheavy = w->nheavy;
  d->nslots = w->nslots;
  if ((rc = dupload(e, (void **)&d->kw.start, w->start,
                    (w->nitems + 1) * sizeof(uint32_t))) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dupload(e, (void **)&d->kw.seg, w->seg,
                    (w->nitems + 1) * sizeof(uint32_t))) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dupload(e, (void **)&d->kw.slot, w->slot,
                    (w->nitems + 1) * sizeof(uint32_t))) != SPLATT_SUCCESS)
    return rc;
  if (w->nheavy) {
    if ((rc = dupload(e, (void **)&d->kw.hseg, w->hseg,
                      w->nheavy * sizeof(uint32_t))) != SPLATT_SUCCESS)
      return rc;
    if ((rc = dupload(e, (void **)&d->kw.hslot, w->hslot,
                      (w->nheavy + 1) * sizeof(uint32_t))) != SPLATT_SUCCESS)
      return rc;
  }
  return SPLATT_SUCCESS;
}

static void free_work(dwork *d)
{
  hipFree((void *)d->kw.start);
  hipFree((void *)d->kw.seg);
  hipFree((void *)d->kw.slot);
  hipFree((void *)d->kw.hseg);
  hipFree((void *)d->kw.hslot);
  memset(d, 0, sizeof(*d));
}

static int is_memo_level(dgroup const *g, uint32_t l)
{
  return g->nim > 0 && l + g->nim >= g->b && l + 1 <= g->b;
}

static void const *gather_factor(adatm_engine const *e, uint32_t mode)
{
  return e->prec == SPLATT_ADATM_BF16ROWS ? e->Abf[mode] : e->A[mode];
}

/* ---- instrumentation ---------------------------------------------------- */
static void ev_begin(adatm_engine *e, int cls, double bytes)
{
  if (!e->instrument || e->nev >= 1024) return;
  hipEventRecord(e->ev[2 * e->nev], e->stream);
  e->ev_class[e->nev] = cls;
  e->ev_bytes[e->nev] = bytes;
}

static void ev_end(adatm_engine *e)
{
  if (!e->instrument || e->nev >= 1024) return;
  hipEventRecord(e->ev[2 * e->nev + 1], e->stream);
  e->nev++;
}

/* algorithmic (compulsory one-touch) bytes of one reduce pass into level
 * l-1 from level l; idx = 4 bytes on the device */
static double bytes_reduce(adatm_engine const *e, dgroup const *g, uint32_t l)
{
  double const z = (double)e->vsz, w = (double)e->wsz, R = (double)e->R;
  double const Fp = (double)g->F[l - 1], Fc = (double)g->F[l];
  double const Iu = (double)e->dims[g->order[l]];
  if (l == e->N - 1) /* leaf: values + leaf ids + parent ptr + factor + out */
    return Fc * (z + 4.0) + (Fp + 1.0) * 4.0 + Iu * R * w + Fp * R * z;
  return (Fp + 1.0) * 4.0 + Fc * 4.0 + Fc * R * z + Iu * R * w + Fp * R * z;
}

static double bytes_scatter(adatm_engine const *e, dgroup const *g, uint32_t d)
{
  double const z = (double)e->vsz, w = (double)e->wsz, R = (double)e->R;
  double b = (double)g->F[d] * R * z + (double)e->dims[g->order[d]] * R * z;
  for (uint32_t l = 0; l < d; ++l)
    b += ((double)g->F[l] + 1.0) * 4.0 + (double)g->F[l + 1] * 4.0 +
         (double)e->dims[g->order[l]] * R * w;
  return b;
}

/* ---- one MTTKRP ----------------------------------------------------------- */
static int run_reduce(adatm_engine *e, dgroup *g, uint32_t l, void const *in,
                      int leaf, void *dst)
{
  ev_begin(e, leaf ? KC_LEAF : KC_LEVEL, bytes_reduce(e, g, l));
  KCHK(adatm_k_seg_reduce(e->stream, e->prec, leaf, &g->red[l].kw,
                          g->ptr[l - 1], g->ids[l], in,
                          gather_factor(e, g->order[l]), dst, e->carry, e->R),
       "seg_reduce");
  ev_end(e);
  return SPLATT_SUCCESS;
}

static int mttkrp_root(adatm_engine *e, dgroup *g)
{
  uint32_t const N = e->N;
  void const *in = g->vals;
  int leaf = 1, pp = 0;
  for (uint32_t l = N - 1; l >= 1; --l) {
    void *dst;
    if (l == 1)
      dst = e->M;
    else if (is_memo_level(g, l - 1))
      dst = g->memo[l - 1];
    else {
      dst = e->tmp[pp];
      pp ^= 1;
    }
    int const rc = run_reduce(e, g, l, in, leaf, dst);
    if (rc != SPLATT_SUCCESS) return rc;
    in = dst;
    leaf = 0;
  }
  return SPLATT_SUCCESS;
}

static int mttkrp_reuse(adatm_engine *e, dgroup *g, uint32_t d)
{
  uint32_t const bd = g->b - d;
  uint32_t const indeg = g->nim < bd ? g->nim : bd;
  uint32_t const L = g->b - indeg;
  void const *in = g->memo[L];
  int pp = 0;
  for (uint32_t l = L; l > d; --l) {
    void *dst = e->tmp[pp];
    pp ^= 1;
    int const rc = run_reduce(e, g, l, in, 0, dst);
    if (rc != SPLATT_SUCCESS) return rc;
    in = dst;
  }
  adatm_kpath path;
  memset(&path, 0, sizeof(path));
  path.d = d;
  for (uint32_t l = 1; l <= d; ++l) path.par[l] = g->par[l];
  for (uint32_t l = 1; l < d; ++l) path.ids[l] = g->ids[l];
  for (uint32_t l = 0; l < d; ++l) path.U[l] = gather_factor(e, g->order[l]);
  ev_begin(e, KC_SCATTER, bytes_scatter(e, g, d));
  KCHK(adatm_k_scatter(e->stream, e->prec, &g->scat[d].kw, g->rowptr[d],
                       g->sperm[d], &path, in, e->M, e->carry, e->R),
       "scatter");
  ev_end(e);
  return SPLATT_SUCCESS;
}

static int mttkrp_dev(adatm_engine *e, uint32_t mode)
{
  dgroup *g = e->g + e->use_grp[mode];
  if (e->use_tag[mode] == 0) return mttkrp_root(e, g);
  return mttkrp_reuse(e, g, e->level_of[e->use_grp[mode]][mode]);
}

/* ---- construction --------------------------------------------------------- */
static int engine_alloc_common(adatm_engine *e)
{
  uint32_t const N = e->N, R = e->R;
  int rc;
  e->maxdim = 0;
  for (uint32_t m = 0; m < N; ++m) {
    if (e->dims[m] > e->maxdim) e->maxdim = e->dims[m];
    if ((rc = dalloc(e, &e->A[m], e->dims[m] * R * e->vsz)) != SPLATT_SUCCESS)
      return rc;
    if (e->prec == SPLATT_ADATM_BF16ROWS &&
        (rc = dalloc(e, &e->Abf[m], e->dims[m] * R * 2)) != SPLATT_SUCCESS)
      return rc;
  }
  if ((rc = dalloc(e, &e->M, e->maxdim * R * e->vsz)) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->G, (size_t)N * R * R * sizeof(double))))
    if (rc != SPLATT_SUCCESS) return rc;
  if ((rc = dalloc(e, (void **)&e->lambda, R * sizeof(double))) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->Vinv, (size_t)R * R * sizeof(double))) !=
      SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->part,
                   adatm_k_part_doubles(R) * sizeof(double))) != SPLATT_SUCCESS)
    return rc;
  e->fits_cap = 4096;
  if ((rc = dalloc(e, (void **)&e->fits, e->fits_cap * sizeof(double))) !=
      SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->scal, 16 * sizeof(double))) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->iter, 4 * sizeof(uint32_t))) != SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, (void **)&e->status, 4 * sizeof(int))) != SPLATT_SUCCESS)
    return rc;
  return SPLATT_SUCCESS;
}

/* deepest level that a reuse mode of this group scatters from */
static uint32_t group_pmax(adatm_engine const *e, uint32_t gi)
{
  uint32_t pm = 0;
  for (uint32_t m = 0; m < e->N; ++m)
    if (e->use_grp[m] == gi && e->use_tag[m] == 1) {
      uint32_t const d = e->level_of[gi][m];
      if (d > pm) pm = d;
    }
  return pm;
}

static int engine_add_group(adatm_engine *e, uint32_t gi, adatm_hcsf const *h)
{
  uint32_t const N = e->N, R = e->R;
  dgroup *g = e->g + gi;
  adatm_group const *pg = e->plan.grp + gi;
  int rc;
  g->memo_mode = pg->memo_mode;
  g->b = pg->begin_imten;
  g->nim = pg->n_imten;
  memcpy(g->order, h->order, N * sizeof(uint32_t));
  memcpy(g->F, h->F, N * sizeof(uint64_t));
  /* values in the path precision */
  if (e->vsz == 8) {
    if ((rc = dupload(e, &g->vals, h->vals, h->nnz * 8)) != SPLATT_SUCCESS)
      return rc;
  } else {
    float *tmp = (float *)malloc(h->nnz * sizeof(float));
    if (!tmp) return SPLATT_ERROR_NOMEMORY;
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)h->nnz; ++j) tmp[j] = (float)h->vals[j];
    rc = dupload(e, &g->vals, tmp, h->nnz * sizeof(float));
    free(tmp);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  for (uint32_t l = 1; l < N; ++l)
    if ((rc = dupload(e, (void **)&g->ids[l], h->fids[l],
                      h->F[l] * sizeof(uint32_t))) != SPLATT_SUCCESS)
      return rc;
  for (uint32_t l = 0; l + 1 < N; ++l)
    if ((rc = dupload(e, (void **)&g->ptr[l], h->fptr[l],
                      (h->F[l] + 1) * sizeof(uint32_t))) != SPLATT_SUCCESS)
      return rc;
  g->pmax = group_pmax(e, gi);
  for (uint32_t l = 1; l <= g->pmax; ++l)
    if ((rc = dupload(e, (void **)&g->par[l], h->par[l],
                      h->F[l] * sizeof(uint32_t))) != SPLATT_SUCCESS)
      return rc;
  /* reduction work per level */
  for (uint32_t l = 1; l < N; ++l) {
    adatm_segwork w;
    if ((rc = adatm_segwork_build(h->fptr[l - 1], h->F[l - 1], RED_TILE, &w)) !=
        SPLATT_SUCCESS)
      return rc;
    rc = upload_work(e, &w, &g->red[l]);
    size_t const cb = (size_t)w.nslots * R * e->vsz;
    if (cb > e->carry_bytes) e->carry_bytes = cb;
    adatm_segwork_free(&w);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  /* memo levels */
  for (uint32_t l = 1; l + 1 < N; ++l)
    if (is_memo_level(g, l) &&
        (rc = dalloc(e, &g->memo[l], g->F[l] * R * e->vsz)) != SPLATT_SUCCESS)
      return rc;
  /* temporaries: every non-root level may be a temporary */
  for (uint32_t l = 1; l + 1 < N; ++l) {
    size_t const tb = (size_t)g->F[l] * R * e->vsz;
    if (tb > e->tmp_bytes) e->tmp_bytes = tb;
  }
  /* reuse scatters */
  for (uint32_t m = 0; m < N; ++m) {
    if (e->use_grp[m] != gi || e->use_tag[m] != 1) continue;
    uint32_t const d = e->level_of[gi][m];
    uint64_t const nrows = e->dims[g->order[d]];
    uint32_t *perm = (uint32_t *)malloc(g->F[d] * sizeof(uint32_t));
    uint32_t *rowptr = (uint32_t *)malloc((nrows + 1) * sizeof(uint32_t));
    if (!perm || !rowptr) {
      free(perm);
      free(rowptr);
      return SPLATT_ERROR_NOMEMORY;
    }
    rc = adatm_sort_by_id(h->fids[d], g->F[d], nrows, perm, rowptr);
    adatm_segwork w;
    memset(&w, 0, sizeof(w));
    if (rc == SPLATT_SUCCESS) rc = adatm_segwork_build(rowptr, nrows, SCAT_TILE, &w);
    if (rc == SPLATT_SUCCESS)
      rc = dupload(e, (void **)&g->sperm[d], perm, g->F[d] * sizeof(uint32_t));
    if (rc == SPLATT_SUCCESS)
      rc = dupload(e, (void **)&g->rowptr[d], rowptr,
                   (nrows + 1) * sizeof(uint32_t));
    if (rc == SPLATT_SUCCESS) rc = upload_work(e, &w, &g->scat[d]);
    size_t const cb = (size_t)w.nslots * R * e->vsz;
    if (cb > e->carry_bytes) e->carry_bytes = cb;
    adatm_segwork_free(&w);
    free(perm);
    free(rowptr);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  return SPLATT_SUCCESS;
}

static int engine_finish(adatm_engine *e)
{
  int rc;
  if ((rc = dalloc(e, &e->tmp[0], e->tmp_bytes)) != SPLATT_SUCCESS) return rc;
  if ((rc = dalloc(e, &e->tmp[1], e->tmp_bytes)) != SPLATT_SUCCESS) return rc;
  if ((rc = dalloc(e, &e->carry, e->carry_bytes ? e->carry_bytes : 64)) !=
      SPLATT_SUCCESS)
    return rc;
  return SPLATT_SUCCESS;
}

static int engine_begin(adatm_engine *e, uint32_t N, uint64_t const *dims,
                        uint64_t nnz, uint64_t rank, adatm_plan const *plan,
                        int precision, int device)
{
  if (N < 3 || N > MAXN || rank < 1 || rank > 64 || plan->n_grp < 1 ||
      plan->n_grp > N) {
    adatm_set_error("engine: unsupported shape (N=%u, R=%llu, groups=%u)", N,
                    (unsigned long long)rank, plan->n_grp);
    return SPLATT_ERROR_BADINPUT;
  }
  if (precision < SPLATT_ADATM_FP64 || precision > SPLATT_ADATM_BF16ROWS)
    return SPLATT_ERROR_BADINPUT;
  e->device = device;
  e->N = N;
  e->R = (uint32_t)rank;
  e->prec = precision;
  e->vsz = (precision == SPLATT_ADATM_FP64) ? 8 : 4;
  e->wsz = (precision == SPLATT_ADATM_FP64) ? 8
           : (precision == SPLATT_ADATM_FP32) ? 4 : 2;
  e->nnz = nnz;
  memcpy(e->dims, dims, N * sizeof(uint64_t));
  e->plan = *plan;
  adatm_plan_use(plan, N, e->use_grp, e->use_tag);
  for (uint32_t gi = 0; gi < plan->n_grp; ++gi)
    for (uint32_t l = 0; l < N; ++l) e->level_of[gi][plan->grp[gi].o_opt[l]] = l;
  /* validate the plan: each reuse level must sit above the group's memo */
  for (uint32_t m = 0; m < N; ++m) {
    uint32_t const gi = e->use_grp[m];
    adatm_group const *pg = plan->grp + gi;
    if (e->use_tag[m] == 1) {
      uint32_t const d = e->level_of[gi][m];
      if (pg->n_imten == 0 || d < 1 || d >= pg->begin_imten) {
        adatm_set_error("engine: mode %u cannot reuse group %u's memo", m, gi);
        return SPLATT_ERROR_BADINPUT;
      }
    } else if (e->level_of[gi][m] != 0) {
      adatm_set_error("engine: root mode %u is not at level 0", m);
      return SPLATT_ERROR_BADINPUT;
    }
    if (pg->begin_imten < 1 || pg->begin_imten > N - 1) {
      adatm_set_error("engine: group %u begins its memo at level %u", gi,
                      pg->begin_imten);
      return SPLATT_ERROR_BADINPUT;
    }
    if (pg->n_imten > 0 && pg->n_imten > pg->begin_imten - 1) {
      adatm_set_error("engine: group %u stores %u levels from %u", gi,
                      pg->n_imten, pg->begin_imten);
      return SPLATT_ERROR_BADINPUT;
    }
  }
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking),
         "hipStreamCreate");
  for (int i = 0; i < 2048; ++i)
    HIPCHK(hipEventCreate(&e->ev[i]), "hipEventCreate");
  return engine_alloc_common(e);
}

int adatm_engine_create_hcsf(adatm_hcsf const *groups, uint32_t N,
                             uint64_t const *dims, uint64_t nnz,
                             double ttnormsq, uint64_t rank,
                             adatm_plan const *plan, int precision, int device,
                             adatm_engine **out)
{
  adatm_engine *e = (adatm_engine *)calloc(1, sizeof(adatm_engine));
  if (!e) return SPLATT_ERROR_NOMEMORY;
  int rc = engine_begin(e, N, dims, nnz, rank, plan, precision, device);
  for (uint32_t gi = 0; rc == SPLATT_SUCCESS && gi < plan->n_grp; ++gi)
    rc = engine_add_group(e, gi, groups + gi);
  if (rc == SPLATT_SUCCESS) rc = engine_finish(e);
  e->ttnormsq = ttnormsq;
  if (rc != SPLATT_SUCCESS) {
    char msg[512];
    snprintf(msg, sizeof(msg), "%s", adatm_last_error());
    adatm_engine_destroy(e);
    adatm_set_error("%s", msg);
    return rc;
  }
  *out = e;
  return SPLATT_SUCCESS;
}

int adatm_engine_create(adatm_coo const *t, uint64_t rank,
                        adatm_plan const *plan, int precision, int device,
                        adatm_engine **out)
{
  *out = NULL;
  uint32_t const N = t->nmodes;
  if (N < 3 || N > MAXN || t->nnz == 0) return SPLATT_ERROR_BADINPUT;
  adatm_hcsf *h = (adatm_hcsf *)calloc(plan->n_grp, sizeof(adatm_hcsf));
  if (!h) return SPLATT_ERROR_NOMEMORY;
  int rc = SPLATT_SUCCESS;
  for (uint32_t gi = 0; rc == SPLATT_SUCCESS && gi < plan->n_grp; ++gi)
    rc = adatm_hcsf_build(t, plan->grp[gi].o_opt, h + gi);
  double ss = 0.0;
  #pragma omp parallel for reduction(+ : ss)
  for (int64_t j = 0; j < (int64_t)t->nnz; ++j) ss += t->vals[j] * t->vals[j];
  if (rc == SPLATT_SUCCESS)
    rc = adatm_engine_create_hcsf(h, N, t->dims, t->nnz, ss, rank, plan,
                                  precision, device, out);
  for (uint32_t gi = 0; gi < plan->n_grp; ++gi) adatm_hcsf_free(h + gi);
  free(h);
  return rc;
}

void adatm_engine_destroy(adatm_engine *e)
{
  if (!e) return;
  if (e->stream) hipStreamSynchronize(e->stream);
  if (e->gvalid) {
    hipGraphExecDestroy(e->gexec);
    hipGraphDestroy(e->graph);
  }
  for (uint32_t gi = 0; gi < MAXN; ++gi) {
    dgroup *g = e->g + gi;
    hipFree(g->vals);
    for (uint32_t l = 0; l < MAXN; ++l) {
      hipFree(g->ids[l]);
      hipFree(g->ptr[l]);
      hipFree(g->par[l]);
      hipFree(g->memo[l]);
      hipFree(g->sperm[l]);
      hipFree(g->rowptr[l]);
      free_work(&g->red[l]);
      free_work(&g->scat[l]);
    }
  }
  for (uint32_t m = 0; m < MAXN; ++m) {
    hipFree(e->A[m]);
    hipFree(e->Abf[m]);
  }
  hipFree(e->M);
  hipFree(e->G);
  hipFree(e->lambda);
  hipFree(e->Vinv);
  hipFree(e->part);
  hipFree(e->fits);
  hipFree(e->scal);
  hipFree(e->iter);
  hipFree(e->status);
  hipFree(e->tmp[0]);
  hipFree(e->tmp[1]);
  hipFree(e->carry);
  if (e->ev[0])
    for (int i = 0; i < 2048; ++i)
      if (e->ev[i]) hipEventDestroy(e->ev[i]);
  if (e->stream) hipStreamDestroy(e->stream);
  free(e);
}

/* ---- factors ------------------------------------------------------------ */
static int factor_from_host(adatm_engine *e, uint32_t mode, double const *A)
{
  size_t const n = e->dims[mode] * e->R;
  double *d64 = NULL;
  int rc = dupload(e, (void **)&d64, A, n * sizeof(double));
  if (rc != SPLATT_SUCCESS) return rc;
  int const k = adatm_k_convert_in(e->stream, e->prec, d64, e->A[mode],
                                   e->Abf[mode], n);
  hipError_t const s = hipStreamSynchronize(e->stream);
  hipFree(d64);
  e->dev_bytes -= n * sizeof(double);
  if (k != 0) return hip_fail((hipError_t)k, "convert_in");
  HIPCHK(s, "convert_in sync");
  e->gvalid = e->gvalid; /* graphs read factors through pointers: still valid */
  return SPLATT_SUCCESS;
}

int adatm_engine_set_factor(adatm_engine *e, uint32_t mode, double const *A)
{
  if (!e || mode >= e->N || !A) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  return factor_from_host(e, mode, A);
}

int adatm_engine_get_factor(adatm_engine *e, uint32_t mode, double *A)
{
  if (!e || mode >= e->N || !A) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  size_t const n = e->dims[mode] * e->R;
  double *d64 = NULL;
  int rc = dalloc(e, (void **)&d64, n * sizeof(double));
  if (rc != SPLATT_SUCCESS) return rc;
  KCHK(adatm_k_convert_out(e->stream, e->prec, e->A[mode], d64, n),
       "convert_out");
  hipError_t const r = hipMemcpyAsync(A, d64, n * sizeof(double),
                                      hipMemcpyDeviceToHost, e->stream);
  hipError_t const s = hipStreamSynchronize(e->stream);
  hipFree(d64);
  e->dev_bytes -= n * sizeof(double);
  HIPCHK(r, "download factor");
  HIPCHK(s, "download sync");
  return SPLATT_SUCCESS;
}

int adatm_engine_get_lambda(adatm_engine *e, double *lambda)
{
  if (!e || !lambda) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(e->stream), "sync");
  HIPCHK(hipMemcpy(lambda, e->lambda, e->R * sizeof(double),
                   hipMemcpyDeviceToHost),
         "download lambda");
  return SPLATT_SUCCESS;
}

int adatm_engine_init_factors(adatm_engine *e, uint64_t seed)
{
  if (!e) return SPLATT_ERROR_BADINPUT;
  for (uint32_t m = 0; m < e->N; ++m) {
    size_t const n = e->dims[m] * e->R;
    double *h = (double *)malloc(n * sizeof(double));
    if (!h) return SPLATT_ERROR_NOMEMORY;
    adatm_rand_matrix(seed, m, e->dims[m], e->R, h);
    int const rc = adatm_engine_set_factor(e, m, h);
    free(h);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  return SPLATT_SUCCESS;
}

/* ---- MTTKRP entry points ------------------------------------------------ */
int adatm_engine_mttkrp(adatm_engine *e, uint32_t mode, double *out)
{
  if (!e || mode >= e->N) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int rc = mttkrp_dev(e, mode);
  if (rc != SPLATT_SUCCESS) return rc;
  if (out) {
    size_t const n = e->dims[mode] * e->R;
    double *d64 = NULL;
    rc = dalloc(e, (void **)&d64, n * sizeof(double));
    if (rc != SPLATT_SUCCESS) return rc;
    int const k = adatm_k_convert_out(e->stream, e->prec, e->M, d64, n);
    hipError_t const r = hipMemcpyAsync(out, d64, n * sizeof(double),
                                        hipMemcpyDeviceToHost, e->stream);
    hipError_t const s = hipStreamSynchronize(e->stream);
    hipFree(d64);
    e->dev_bytes -= n * sizeof(double);
    if (k != 0) return hip_fail((hipError_t)k, "convert_out");
    HIPCHK(r, "download M");
    HIPCHK(s, "mttkrp sync");
  } else {
    HIPCHK(hipStreamSynchronize(e->stream), "mttkrp sync");
  }
  return SPLATT_SUCCESS;
}

static int seq_enqueue(adatm_engine *e)
{
  for (uint32_t m = 0; m < e->N; ++m) {
    int const rc = mttkrp_dev(e, m);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  return SPLATT_SUCCESS;
}

int adatm_engine_mttkrp_seq(adatm_engine *e)
{
  if (!e) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int const rc = seq_enqueue(e);
  if (rc != SPLATT_SUCCESS) return rc;
  HIPCHK(hipStreamSynchronize(e->stream), "seq sync");
  return SPLATT_SUCCESS;
}

/* ---- CP-ALS --------------------------------------------------------------- */
static int sweep_enqueue(adatm_engine *e)
{
  uint32_t const N = e->N, R = e->R;
  for (uint32_t m = 0; m < N; ++m) {
    int const rc = mttkrp_dev(e, m);
    if (rc != SPLATT_SUCCESS) return rc;
    ev_begin(e, KC_DENSE, 0.0);
    KCHK(adatm_k_solve_inv(e->stream, e->G, N, m, R, e->Vinv, e->status),
         "solve_inv");
    KCHK(adatm_k_update(e->stream, e->prec, e->M, e->Vinv, e->A[m], e->Abf[m],
                        e->dims[m], R, e->iter, e->lambda, e->part,
                        e->G + (size_t)m * R * R),
         "update");
    ev_end(e);
  }
  ev_begin(e, KC_DENSE, 0.0);
  KCHK(adatm_k_fit(e->stream, e->prec, e->A[N - 1], e->M, e->dims[N - 1],
                   e->lambda, e->G, N, R, e->ttnormsq, e->part, e->fits,
                   e->iter),
       "fit");
  ev_end(e);
  return SPLATT_SUCCESS;
}

/* Grams of the current factors and the sweep counter reset */
static int cpd_prepare(adatm_engine *e)
{
  uint32_t const R = e->R;
  for (uint32_t m = 0; m < e->N; ++m)
    KCHK(adatm_k_gram(e->stream, e->prec, e->A[m], e->dims[m], R, e->part,
                      e->G + (size_t)m * R * R),
         "gram");
  KCHK(adatm_k_zero(e->stream, e->iter, 4 * sizeof(uint32_t)), "zero iter");
  return SPLATT_SUCCESS;
}

static int graph_build(adatm_engine *e)
{
  if (e->gvalid) return SPLATT_SUCCESS;
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal),
         "begin capture");
  int const rc = sweep_enqueue(e);
  hipGraph_t g = NULL;
  hipError_t const ec = hipStreamEndCapture(e->stream, &g);
  if (rc != SPLATT_SUCCESS) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  HIPCHK(ec, "end capture");
  hipError_t const ei = hipGraphInstantiate(&e->gexec, g, NULL, NULL, 0);
  if (ei != hipSuccess) {
    hipGraphDestroy(g);
    return hip_fail(ei, "graph instantiate");
  }
  e->graph = g;
  e->gvalid = 1;
  return SPLATT_SUCCESS;
}

static int sweep_launch(adatm_engine *e, int use_graph)
{
  if (use_graph) {
    int const rc = graph_build(e);
    if (rc != SPLATT_SUCCESS) return rc;
    HIPCHK(hipGraphLaunch(e->gexec, e->stream), "graph launch");
    return SPLATT_SUCCESS;
  }
  return sweep_enqueue(e);
}

int adatm_engine_cpd(adatm_engine *e, uint32_t niters, double tol,
                     int use_graph, double *fits, uint32_t *iters_done)
{
  if (!e || niters == 0 || niters > e->fits_cap) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int rc = cpd_prepare(e);
  if (rc != SPLATT_SUCCESS) return rc;
  uint32_t done = 0;
  double oldfit = 0.0;
  for (uint32_t it = 0; it < niters; ++it) {
    rc = sweep_launch(e, use_graph);
    if (rc != SPLATT_SUCCESS) return rc;
    done = it + 1;
    if (tol > 0.0) {
      double fit;
      HIPCHK(hipMemcpyAsync(&fit, e->fits + it, sizeof(double),
                            hipMemcpyDeviceToHost, e->stream),
             "fit download");
      HIPCHK(hipStreamSynchronize(e->stream), "sweep sync");
      if (it > 0 && fabs(fit - oldfit) < tol) break;
      oldfit = fit;
    }
  }
  HIPCHK(hipStreamSynchronize(e->stream), "cpd sync");
  if (fits)
    HIPCHK(hipMemcpy(fits, e->fits, done * sizeof(double),
                     hipMemcpyDeviceToHost),
           "fits download");
  if (iters_done) *iters_done = done;
  return SPLATT_SUCCESS;
}

/* ---- timing ----------------------------------------------------------------- */
static int cmp_double(void const *a, void const *b)
{
  double const x = *(double const *)a, y = *(double const *)b;
  return (x > y) - (x < y);
}

int adatm_engine_time(adatm_engine *e, int what, int nwarm, int nrep,
                      int use_graph, double *ms_median, double *per_kernel_ms)
{
  if (!e || nrep < 1 || nrep > 1000) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int rc;
  if (what == 1) {
    rc = cpd_prepare(e);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  /* sweeps write fits[iter]: keep the counter inside the buffer */
  for (int i = 0; i < nwarm; ++i) {
    if (what == 1 && (uint32_t)(i + 1) >= e->fits_cap) break;
    rc = (what == 0) ? seq_enqueue(e) : sweep_launch(e, use_graph);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  HIPCHK(hipStreamSynchronize(e->stream), "warmup sync");
  double *t = (double *)malloc((size_t)nrep * sizeof(double));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a), "event");
  HIPCHK(hipEventCreate(&b), "event");
  for (int i = 0; i < nrep; ++i) {
    if (what == 1 && (uint32_t)(nwarm + i + 1) >= e->fits_cap) {
      rc = cpd_prepare(e);
      if (rc != SPLATT_SUCCESS) break;
    }
    hipEventRecord(a, e->stream);
    rc = (what == 0) ? seq_enqueue(e) : sweep_launch(e, use_graph && what == 1);
    if (rc != SPLATT_SUCCESS) break;
    hipEventRecord(b, e->stream);
    hipError_t const s = hipEventSynchronize(b);
    if (s != hipSuccess) {
      rc = hip_fail(s, "timing sync");
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    t[i] = ms;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != SPLATT_SUCCESS) {
    free(t);
    return rc;
  }
  qsort(t, (size_t)nrep, sizeof(double), cmp_double);
  *ms_median = (nrep % 2) ? t[nrep / 2] : 0.5 * (t[nrep / 2 - 1] + t[nrep / 2]);
  free(t);

  if (per_kernel_ms) {
    /* one instrumented, non-graph pass: per-class device time and bytes */
    double tsum[KC_NCLASS] = {0}, bsum[KC_NCLASS] = {0};
    int cnt[KC_NCLASS] = {0};
    e->instrument = 1;
    e->nev = 0;
    rc = (what == 0) ? seq_enqueue(e) : sweep_enqueue(e);
    e->instrument = 0;
    if (rc != SPLATT_SUCCESS) return rc;
    HIPCHK(hipStreamSynchronize(e->stream), "instrument sync");
    for (int i = 0; i < e->nev; ++i) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e->ev[2 * i], e->ev[2 * i + 1]);
      tsum[e->ev_class[i]] += ms;
      bsum[e->ev_class[i]] += e->ev_bytes[i];
      cnt[e->ev_class[i]]++;
    }
    /* layout: [c] total ms, [8+c] launches, [16+c] algorithmic bytes */
    for (int c = 0; c < KC_NCLASS; ++c) {
      per_kernel_ms[c] = tsum[c];
      per_kernel_ms[8 + c] = (double)cnt[c];
      per_kernel_ms[16 + c] = bsum[c];
    }
  }
  return SPLATT_SUCCESS;
}

/* ---- statistics --------------------------------------------------------------- */
int adatm_engine_stats(adatm_engine *e, uint64_t *nfibs, double *seq_bytes,
                       double *seq_flops, uint64_t *device_bytes)
{
  if (!e) return SPLATT_ERROR_BADINPUT;
  uint32_t const N = e->N;
  if (nfibs)
    for (uint32_t gi = 0; gi < e->plan.n_grp; ++gi)
      for (uint32_t l = 0; l < N; ++l) nfibs[gi * N + l] = e->g[gi].F[l];
  double B = 0.0, Fl = 0.0;
  double const R = (double)e->R;
  for (uint32_t m = 0; m < N; ++m) {
    dgroup const *g = e->g + e->use_grp[m];
    if (e->use_tag[m] == 0) {
      for (uint32_t l = N - 1; l >= 1; --l) {
        B += bytes_reduce(e, g, l);
        Fl += 2.0 * R * (double)g->F[l];
      }
    } else {
      uint32_t const d = e->level_of[e->use_grp[m]][m];
      uint32_t const bd = g->b - d;
      uint32_t const L = g->b - (g->nim < bd ? g->nim : bd);
      for (uint32_t l = L; l > d; --l) {
        B += bytes_reduce(e, g, l);
        Fl += 2.0 * R * (double)g->F[l];
      }
      B += bytes_scatter(e, g, d);
      Fl += R * (double)(d + 1) * (double)g->F[d];
    }
  }
  if (seq_bytes) *seq_bytes = B;
  if (seq_flops) *seq_flops = Fl;
  if (device_bytes) *device_bytes = e->dev_bytes;
  return SPLATT_SUCCESS;
}

int adatm_engine_get_csf(adatm_engine *e, uint32_t grp, uint32_t **fptr,
                         uint32_t **fids, double *vals)
{
  if (!e || grp >= e->plan.n_grp) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(e->stream), "sync");
  dgroup const *g = e->g + grp;
  uint32_t const N = e->N;
  for (uint32_t l = 0; l + 1 < N; ++l)
    if (fptr && fptr[l])
      HIPCHK(hipMemcpy(fptr[l], g->ptr[l], (g->F[l] + 1) * sizeof(uint32_t),
                       hipMemcpyDeviceToHost),
             "download fptr");
  for (uint32_t l = 1; l < N; ++l)
    if (fids && fids[l])
      HIPCHK(hipMemcpy(fids[l], g->ids[l], g->F[l] * sizeof(uint32_t),
                       hipMemcpyDeviceToHost),
             "download fids");
  if (vals) {
    double *d64 = NULL;
    int rc = dalloc(e, (void **)&d64, e->nnz * sizeof(double));
    if (rc != SPLATT_SUCCESS) return rc;
    KCHK(adatm_k_convert_out(e->stream, e->prec, g->vals, d64, e->nnz),
         "convert vals");
    hipError_t const r = hipMemcpyAsync(vals, d64, e->nnz * sizeof(double),
                                        hipMemcpyDeviceToHost, e->stream);
    hipError_t const s = hipStreamSynchronize(e->stream);
    hipFree(d64);
    e->dev_bytes -= e->nnz * sizeof(double);
    HIPCHK(r, "download vals");
    HIPCHK(s, "sync");
  }
  return SPLATT_SUCCESS;
}

/* ---- accessors for the SPLATT-compatible layer -------------------------- */
uint32_t adatm_engine_nmodes(adatm_engine const *e) { return e->N; }
uint64_t adatm_engine_dim(adatm_engine const *e, uint32_t m) { return e->dims[m]; }
uint32_t adatm_engine_rank(adatm_engine const *e) { return e->R; }
int adatm_engine_precision(adatm_engine const *e) { return e->prec; }
