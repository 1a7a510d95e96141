/*
 * SPLATT-compatible C API of libsplatt_adatm (include/splatt.h).
 *
 * The reference patches SPLATT v1.1.1, whose public API is kept here with
 * identical signatures (reference include/splatt.h:384-525) and the AdaTM
 * entry points that the reference defines (src/adatm_rcsf.c, src/mttkrp.c,
 * src/adatm_cpd.c).  Each CSF array handed out is registered with the device
 * state that serves it: the plan, the (group, tag) use map and -- created on
 * the first call that knows the rank -- an adatm_engine on the GPU.
 *
 * Allocation schemes (opts[SPLATT_OPTION_CSF_ALLOC]):
 *   ADAPTIVE, ADAPTIVE_TIME_EFFICIENT  one CSF per memoization group of the
 *        analytic (or learned) plan; root modes recompute their group's memo,
 *        the others reuse it (src/adatm_rcsf.c:503-526, src/mttkrp.c:1958);
 *   ONEMODE / TWOMODE / ALLMODE  SPLATT's plain CSF sets: the root of a CSF
 *        by a root pass, any other level from the leaves (tag 2), as
 *        mttkrp_csf does (src/mttkrp.c:1885-1936);
 *   ADAPTIVE_SPACE_EFFICIENT keeps the CSFs of the memoising groups only;
 *   the other groups' modes are computed from their leaves (tag 2).
 *
 * Differences from the reference, on purpose:
 *   - splatt_mttkrp serves adaptive CSFs (the reference's mttkrp_csf has no
 *     adaptive case and leaves the output zero, src/mttkrp.c:1904-1935);
 *   - a reuse MTTKRP checks that the factors its memo contracted are the ones
 *     passed now (content hash) and recomputes the memo when they are not;
 *   - the root level keeps its ids (fids[0] is not NULL), so tensors with
 *     empty slices factor correctly (the untiled reference drops them,
 *     src/adatm_rcsf.c:51-56, and needs every slice non-empty).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "adatm_host.h"

#define MAXN ADATM_MAX_NMODES

typedef struct reg_ent {
  splatt_csf *csf; /* key: the array returned to the caller */
  uint32_t n_csf;
  uint32_t N;
  uint64_t nnz;
  uint64_t dims[MAXN];
  adatm_plan plan;  /* the groups that have a CSF */
  adatm_plan full;  /* the plan as chosen / described by the caller */
  uint32_t use_grp[MAXN], use_tag[MAXN];
  double ttnormsq;
  int device;
  adatm_engine *eng;
  uint32_t eng_rank;
  int eng_prec;
  uint64_t fac_hash[MAXN];
  int fac_on_dev[MAXN];
  uint64_t memo_fac[MAXN][MAXN]; /* per group: factor hashes at its last root */
  int memo_ok[MAXN];
  adatm_coo coo; /* the coordinates as given, kept for device engine builds */
  struct reg_ent *next;
} reg_ent;

static reg_ent *g_reg = NULL;
static pthread_mutex_t g_reg_lock = PTHREAD_MUTEX_INITIALIZER;

static reg_ent *reg_find(splatt_csf const *csf)
{
  pthread_mutex_lock(&g_reg_lock);
  reg_ent *r = g_reg;
  while (r && r->csf != csf) r = r->next;
  pthread_mutex_unlock(&g_reg_lock);
  return r;
}

static void reg_add(reg_ent *r)
{
  pthread_mutex_lock(&g_reg_lock);
  r->next = g_reg;
  g_reg = r;
  pthread_mutex_unlock(&g_reg_lock);
}

static reg_ent *reg_remove(splatt_csf const *csf)
{
  pthread_mutex_lock(&g_reg_lock);
  reg_ent **pp = &g_reg;
  while (*pp && (*pp)->csf != csf) pp = &(*pp)->next;
  reg_ent *r = *pp;
  if (r) *pp = r->next;
  pthread_mutex_unlock(&g_reg_lock);
  return r;
}

/* ---- options (reference include/splatt.h:384-391) ----------------------- */
double *splatt_default_opts(void)
{
  double *o = (double *)malloc(SPLATT_OPTION_NOPTIONS * sizeof(double));
  if (!o) return NULL;
#ifdef _OPENMP
  o[SPLATT_OPTION_NTHREADS] = omp_get_max_threads();
#else
  o[SPLATT_OPTION_NTHREADS] = 1;
#endif
  o[SPLATT_OPTION_TOLERANCE] = 1e-5;
  o[SPLATT_OPTION_NITER] = 50;
  o[SPLATT_OPTION_VERBOSITY] = SPLATT_VERBOSITY_LOW;
  o[SPLATT_OPTION_RANDSEED] = (double)time(NULL);
  o[SPLATT_OPTION_CSF_ALLOC] = SPLATT_ADAPTIVE_TIME_EFFICIENT;
  o[SPLATT_OPTION_TILE] = SPLATT_NOTILE;
  o[SPLATT_OPTION_TILEDEPTH] = 1;
  o[SPLATT_OPTION_DECOMP] = SPLATT_DECOMP_MEDIUM;
  o[SPLATT_OPTION_COMM] = SPLATT_COMM_ALL2ALL;
  o[SPLATT_OPTION_ADATM_PRECISION] = SPLATT_ADATM_FP64;
  o[SPLATT_OPTION_ADATM_PLANNER] = SPLATT_ADATM_PLAN_ANALYTIC;
  o[SPLATT_OPTION_ADATM_GRAPH] = 1;
  o[SPLATT_OPTION_ADATM_REFQUIRKS] = 1;
  o[SPLATT_OPTION_ADATM_SPACE_LIMIT] = (double)SPLATT_ADATM_SLIMIT;
  o[SPLATT_OPTION_ADATM_DEVICE] = 0;
  o[SPLATT_OPTION_ADATM_RANK_HINT] = 16;
  return o;
}

void splatt_free_opts(double *opts) { free(opts); }

/* ---- plain SPLATT CSF sets as plans ------------------------------------- */
/* modes sorted by dimension, shortest first, ties by mode index */
static void small_first(uint32_t N, uint64_t const *dims, uint32_t *order)
{
  for (uint32_t m = 0; m < N; ++m) order[m] = m;
  for (uint32_t i = 1; i < N; ++i)
    for (uint32_t j = i; j > 0 && dims[order[j]] < dims[order[j - 1]]; --j) {
      uint32_t const t = order[j];
      order[j] = order[j - 1];
      order[j - 1] = t;
    }
}

static void root_first(uint32_t N, uint64_t const *dims, uint32_t root,
                       uint32_t *order)
{
  uint32_t sf[MAXN], k = 1;
  small_first(N, dims, sf);
  order[0] = root;
  for (uint32_t i = 0; i < N; ++i)
    if (sf[i] != root) order[k++] = sf[i];
}

static int plain_plan(int alloc, uint32_t N, uint64_t const *dims,
                      adatm_plan *p, uint32_t *use_grp, uint32_t *use_tag)
{
  memset(p, 0, sizeof(*p));
  uint32_t o[MAXN] = {0};
  small_first(N, dims, o);
  switch (alloc) {
  case SPLATT_CSF_ONEMODE:
    p->n_grp = 1;
    memcpy(p->grp[0].o_opt, o, N * sizeof(uint32_t));
    p->grp[0].memo_mode = o[0];
    p->grp[0].begin_imten = 1;
    for (uint32_t m = 0; m < N; ++m) {
      use_grp[m] = 0;
      use_tag[m] = (m == o[0]) ? 0 : 2;
    }
    return SPLATT_SUCCESS;
  case SPLATT_CSF_TWOMODE:
    p->n_grp = 2;
    memcpy(p->grp[0].o_opt, o, N * sizeof(uint32_t));
    p->grp[0].memo_mode = o[0];
    p->grp[0].begin_imten = 1;
    root_first(N, dims, o[N - 1], p->grp[1].o_opt);
    p->grp[1].memo_mode = o[N - 1];
    p->grp[1].begin_imten = 1;
    for (uint32_t m = 0; m < N; ++m) {
      if (m == o[N - 1]) {
        use_grp[m] = 1;
        use_tag[m] = 0;
      } else {
        use_grp[m] = 0;
        use_tag[m] = (m == o[0]) ? 0 : 2;
      }
    }
    return SPLATT_SUCCESS;
  case SPLATT_CSF_ALLMODE:
    p->n_grp = N;
    for (uint32_t m = 0; m < N; ++m) {
      root_first(N, dims, m, p->grp[m].o_opt);
      p->grp[m].memo_mode = m;
      p->grp[m].begin_imten = 1;
      use_grp[m] = m;
      use_tag[m] = 0;
    }
    return SPLATT_SUCCESS;
  default:
    return SPLATT_ERROR_BADINPUT;
  }
}

/* ---- host CSF (u32) <-> SPLATT layout ------------------------------------ */
/* one tile's arrays from the subtree of levels lv0.. of h rooted at the
 * level-lv0 nodes [lo, hi) (lv0 = 0, whole range: the untiled CSF) */
static int fill_sparsity(csf_sparsity *pt, adatm_hcsf const *h, uint32_t lv0, uint64_t lo,
                         uint64_t hi)
{
  uint32_t const N = h->N - lv0;
  for (uint32_t l = 0; l < N; ++l) {
    uint32_t const hl = lv0 + l;
    uint64_t const F = hi - lo;
    pt->nfibs[l] = F;
    pt->fids[l] = (splatt_idx_t *)malloc((F ? F : 1) * sizeof(splatt_idx_t));
    if (!pt->fids[l]) return SPLATT_ERROR_NOMEMORY;
    for (uint64_t i = 0; i < F; ++i) pt->fids[l][i] = h->fids[hl][lo + i];
    if (l + 1 < N) {
      uint32_t const *fp = h->fptr[hl];
      pt->fptr[l] = (splatt_idx_t *)malloc((F + 1) * sizeof(splatt_idx_t));
      if (!pt->fptr[l]) return SPLATT_ERROR_NOMEMORY;
      for (uint64_t i = 0; i <= F; ++i) pt->fptr[l][i] = fp[lo + i] - fp[lo];
      uint64_t const nlo = fp[lo], nhi = fp[hi];
      lo = nlo;
      hi = nhi;
    } else {
      pt->vals = (splatt_val_t *)malloc((F ? F : 1) * sizeof(splatt_val_t));
      if (!pt->vals) return SPLATT_ERROR_NOMEMORY;
      memcpy(pt->vals, h->vals + lo, F * sizeof(double));
    }
  }
  return SPLATT_SUCCESS;
}

static void csf_head(splatt_csf *c, uint32_t N, uint64_t nnz, uint64_t const *dims,
                     uint32_t const *order)
{
  c->nnz = nnz;
  c->nmodes = N;
  for (uint32_t m = 0; m < N; ++m) {
    c->dims[m] = dims[m];
    c->dim_perm[m] = order[m];
    c->tile_dims[m] = 1;
  }
  c->which_tile = SPLATT_NOTILE;
  c->ntiles = 1;
}

static int csf_from_hcsf(splatt_csf *c, adatm_hcsf const *h, uint64_t const *dims)
{
  csf_head(c, h->N, h->nnz, dims, h->order);
  c->pt = (csf_sparsity *)calloc(1, sizeof(csf_sparsity));
  if (!c->pt) return SPLATT_ERROR_NOMEMORY;
  return fill_sparsity(c->pt, h, 0, 0, h->F[0]);
}

/*
 * SPLATT's dense tiling (SPLATT_DENSETILE; p_csf_alloc_densetile and
 * tt_densetile of SPLATT v1.1.1, selected by --tile at cmd_bench.c:151 and
 * cmd_cpd.c:90).  Every level at depth >= `depth` is cut into `ntile`
 * equal index ranges (tile size max(dim / ntile, 1), the last range takes
 * the rest), the others into one; tile ids are row-major over the modes
 * (the last mode fastest) and each tile holds the CSF of its nonzeros in
 * the untiled CSF order.  Built on the device as one CSF with the tile id
 * as an extra leading level, then cut per tile.  The MTTKRP of a tiled set
 * runs on the GPU's own (untiled) structures: tiling only schedules CPU
 * threads and does not change the result beyond summation order.
 */
#define MAX_TILES (1u << 20)
static int csf_densetile(splatt_csf *c, adatm_coo const *t, uint32_t const *order,
                         uint64_t ntile, uint64_t depth, int device)
{
  uint32_t const N = t->nmodes;
  csf_head(c, N, t->nnz, t->dims, order);
  uint64_t td[MAXN], tsz[MAXN], ntiles = 1;
  for (uint32_t m = 0; m < N; ++m) td[m] = 1;
  for (uint32_t d = (uint32_t)depth; d < N; ++d) td[order[d]] = ntile ? ntile : 1;
  for (uint32_t m = 0; m < N; ++m) {
    ntiles *= td[m];
    if (ntiles > MAX_TILES) {
      adatm_set_error("dense tiling: more than %u tiles", MAX_TILES);
      return SPLATT_ERROR_BADINPUT;
    }
    tsz[m] = t->dims[m] / td[m] ? t->dims[m] / td[m] : 1;
    c->tile_dims[m] = td[m];
  }
  c->which_tile = SPLATT_DENSETILE;
  c->ntiles = ntiles;
  c->pt = (csf_sparsity *)calloc(ntiles, sizeof(csf_sparsity));
  if (!c->pt) return SPLATT_ERROR_NOMEMORY;
  /* the tile id as mode N, sorted first */
  adatm_coo x = *t;
  x.nmodes = N + 1;
  x.dims[N] = ntiles;
  x.ind[N] = (uint32_t *)malloc(t->nnz * sizeof(uint32_t));
  if (!x.ind[N]) return SPLATT_ERROR_NOMEMORY;
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)t->nnz; ++j) {
    uint64_t id = 0;
    for (uint32_t m = 0; m < N; ++m) {
      uint64_t cm = t->ind[m][j] / tsz[m];
      if (cm >= td[m]) cm = td[m] - 1;
      id = id * td[m] + cm;
    }
    x.ind[N][j] = (uint32_t)id;
  }
  uint32_t xo[MAXN + 1];
  xo[0] = N;
  memcpy(xo + 1, order, N * sizeof(uint32_t));
  adatm_hcsf h;
  int rc = device >= 0 ? adatm_hcsf_build_device(&x, xo, device, &h)
                       : adatm_hcsf_build(&x, xo, &h);
  free(x.ind[N]);
  if (rc != SPLATT_SUCCESS) return rc;
  for (uint64_t k = 0; rc == SPLATT_SUCCESS && k < h.F[0]; ++k)
    rc = fill_sparsity(c->pt + h.fids[0][k], &h, 1, h.fptr[0][k], h.fptr[0][k + 1]);
  /* empty tiles: no fibers, a two-entry root pointer (SPLATT's layout) */
  for (uint64_t tl = 0; rc == SPLATT_SUCCESS && tl < ntiles; ++tl)
    if (!c->pt[tl].fids[0]) {
      c->pt[tl].fptr[0] = (splatt_idx_t *)calloc(2, sizeof(splatt_idx_t));
      if (!c->pt[tl].fptr[0]) rc = SPLATT_ERROR_NOMEMORY;
    }
  adatm_hcsf_free(&h);
  return rc;
}

static void free_csf_array(splatt_csf *t, uint32_t n)
{
  if (!t) return;
  for (uint32_t g = 0; g < n; ++g) {
    if (!t[g].pt) continue;
    for (uint64_t k = 0; k < t[g].ntiles; ++k) {
      csf_sparsity *pt = t[g].pt + k;
      for (uint32_t l = 0; l < SPLATT_MAX_NMODES; ++l) {
        free(pt->fptr[l]);
        free(pt->fids[l]);
      }
      free(pt->vals);
    }
    free(t[g].pt);
  }
  free(t);
}

/* ---- construction ---------------------------------------------------------- */
static int build_from_coo(adatm_coo const *t, double const *opts,
                          adatm_plan const *given, adatm_plan const *given_full,
                          uint32_t const *given_grp, uint32_t const *given_tag,
                          splatt_csf **tensors, uint32_t *n_csf_out)
{
  uint32_t const N = t->nmodes;
  int const alloc = (int)opts[SPLATT_OPTION_CSF_ALLOC];
  reg_ent *r = (reg_ent *)calloc(1, sizeof(reg_ent));
  if (!r) return SPLATT_ERROR_NOMEMORY;
  r->N = N;
  r->nnz = t->nnz;
  memcpy(r->dims, t->dims, N * sizeof(uint64_t));
  r->device = (int)opts[SPLATT_OPTION_ADATM_DEVICE];
  int rc = SPLATT_SUCCESS;
  if (given) {
    r->plan = *given;
    r->full = given_full ? *given_full : *given;
    memcpy(r->use_grp, given_grp, N * sizeof(uint32_t));
    memcpy(r->use_tag, given_tag, N * sizeof(uint32_t));
  } else if (alloc == SPLATT_ADAPTIVE || alloc == SPLATT_ADAPTIVE_TIME_EFFICIENT) {
    adatm_plan cands[256];
    int nc = 0, ch = 0;
    int const planner = (int)opts[SPLATT_OPTION_ADATM_PLANNER];
    uint64_t const hint = opts[SPLATT_OPTION_ADATM_RANK_HINT] >= 1
                              ? (uint64_t)opts[SPLATT_OPTION_ADATM_RANK_HINT]
                              : 16;
    if (planner == SPLATT_ADATM_PLAN_LEARNED) {
      adatm_mlp *model = NULL;
      char path[4096];
      if (adatm_costmodel_path(path, sizeof(path)) &&
          adatm_mlp_load(r->device, path, &model) != SPLATT_SUCCESS)
        model = NULL; /* no usable model: the analytic choice */
      rc = adatm_plan_learned(t, hint, model, r->device, cands, 256, &nc, &ch, NULL);
      adatm_mlp_destroy(model);
    } else if (planner == SPLATT_ADATM_PLAN_EXHAUSTIVE) {
      int const prec = (int)opts[SPLATT_OPTION_ADATM_PRECISION];
      rc = adatm_plan_exhaustive(t, hint, prec, r->device, 1, 3, cands, 256, &nc, &ch,
                                 NULL);
    } else {
      /* the rank of the decomposition is not known yet: plan for the hint
         (the analytic choice, min pops, is rank-independent; only the space
         filter depends on the rank); fiber counts on the device */
      rc = adatm_plan_analytic(t, hint, (int)opts[SPLATT_OPTION_ADATM_REFQUIRKS],
                               opts[SPLATT_OPTION_ADATM_SPACE_LIMIT], r->device, cands,
                               256, &nc, &ch);
    }
    if (rc == SPLATT_SUCCESS) {
      r->plan = cands[ch];
      r->full = r->plan;
      adatm_plan_use(&r->plan, N, r->use_grp, r->use_tag);
    }
  } else if (alloc == SPLATT_ADAPTIVE_SPACE_EFFICIENT) {
    adatm_plan cands[256], full;
    int nc = 0, ch = 0;
    uint64_t const hint = opts[SPLATT_OPTION_ADATM_RANK_HINT] >= 1
                              ? (uint64_t)opts[SPLATT_OPTION_ADATM_RANK_HINT]
                              : 16;
    rc = adatm_plan_analytic(t, hint, (int)opts[SPLATT_OPTION_ADATM_REFQUIRKS],
                             opts[SPLATT_OPTION_ADATM_SPACE_LIMIT], r->device, cands, 256,
                             &nc, &ch);
    if (rc == SPLATT_SUCCESS) {
      full = cands[ch];
      r->full = full;
      adatm_plan_space_efficient(&full, N, &r->plan, r->use_grp, r->use_tag);
    }
  } else {
    rc = plain_plan(alloc, N, t->dims, &r->plan, r->use_grp, r->use_tag);
    r->full = r->plan;
  }
  if (rc != SPLATT_SUCCESS) {
    free(r);
    return rc;
  }
  uint32_t const ng = r->plan.n_grp;
  /* SPLATT's CSF sets may be tiled; AdaTM's never are (adatm_rcsf.c:304) */
  int const tiled = !given && (int)opts[SPLATT_OPTION_TILE] == SPLATT_DENSETILE &&
                    (alloc == SPLATT_CSF_ONEMODE || alloc == SPLATT_CSF_TWOMODE ||
                     alloc == SPLATT_CSF_ALLMODE);
  splatt_csf *arr = (splatt_csf *)calloc(ng, sizeof(splatt_csf));
  if (!arr) rc = SPLATT_ERROR_NOMEMORY;
  /* on the device: the GPU radix sort and scans; the host build otherwise */
  for (uint32_t g = 0; rc == SPLATT_SUCCESS && g < ng; ++g) {
    if (tiled) {
      double const nt = opts[SPLATT_OPTION_NTHREADS], dp = opts[SPLATT_OPTION_TILEDEPTH];
      rc = csf_densetile(arr + g, t, r->plan.grp[g].o_opt, nt >= 1 ? (uint64_t)nt : 1,
                         dp >= 0 ? (uint64_t)dp : 0, r->device);
      continue;
    }
    adatm_hcsf h;
    rc = (r->device >= 0) ? adatm_hcsf_build_device(t, r->plan.grp[g].o_opt, r->device, &h)
                          : adatm_hcsf_build(t, r->plan.grp[g].o_opt, &h);
    if (rc == SPLATT_SUCCESS) rc = csf_from_hcsf(arr + g, &h, t->dims);
    adatm_hcsf_free(&h);
  }
  if (rc == SPLATT_SUCCESS && r->device >= 0) {
    /* keep the coordinates: engines for this CSF set are built on the device */
    rc = adatm_coo_alloc(N, t->nnz, t->dims, &r->coo);
    for (uint32_t m = 0; rc == SPLATT_SUCCESS && m < N; ++m)
      memcpy(r->coo.ind[m], t->ind[m], t->nnz * sizeof(uint32_t));
    if (rc == SPLATT_SUCCESS) memcpy(r->coo.vals, t->vals, t->nnz * sizeof(double));
  }
  if (rc != SPLATT_SUCCESS) {
    free_csf_array(arr, ng);
    adatm_coo_free(&r->coo);
    free(r);
    return rc;
  }
  r->ttnormsq = adatm_sumsq(t->vals, t->nnz);
  r->csf = arr;
  r->n_csf = ng;
  reg_add(r);
  *tensors = arr;
  if (n_csf_out) *n_csf_out = ng;
  return SPLATT_SUCCESS;
}

int splatt_csf_convert(splatt_idx_t const nmodes, splatt_idx_t const nnz,
                       splatt_idx_t **const inds, splatt_val_t *const vals,
                       splatt_csf **tensors, double const *const options)
{
  if (!tensors || !inds || !vals || nmodes < 3 || nmodes > MAXN || nnz == 0 ||
      nnz >= (1ull << 32)) {
    adatm_set_error("splatt_csf_convert: unsupported input (nmodes=%llu, nnz=%llu)",
                    (unsigned long long)nmodes, (unsigned long long)nnz);
    return SPLATT_ERROR_BADINPUT;
  }
  double *dflt = NULL;
  double const *opts = options;
  if (!opts) opts = dflt = splatt_default_opts();
  adatm_coo t;
  uint64_t dims[MAXN];
  for (uint32_t m = 0; m < nmodes; ++m) {
    uint64_t mx = 0;
    for (uint64_t j = 0; j < nnz; ++j)
      if (inds[m][j] > mx) mx = inds[m][j];
    dims[m] = mx + 1;
    if (dims[m] > (1ull << 31)) {
      free(dflt);
      return SPLATT_ERROR_BADINPUT;
    }
  }
  int rc = adatm_coo_alloc((uint32_t)nmodes, nnz, dims, &t);
  if (rc != SPLATT_SUCCESS) {
    free(dflt);
    return rc;
  }
  for (uint32_t m = 0; m < nmodes; ++m)
    for (uint64_t j = 0; j < nnz; ++j) t.ind[m][j] = (uint32_t)inds[m][j];
  memcpy(t.vals, vals, nnz * sizeof(double));
  rc = build_from_coo(&t, opts, NULL, NULL, NULL, NULL, tensors, NULL);
  adatm_coo_free(&t);
  free(dflt);
  return rc;
}

splatt_csf *splatt_csf_alloc(sptensor_t *const tt, double const *const opts)
{
  splatt_csf *out = NULL;
  if (!tt) return NULL;
  if (splatt_csf_convert(tt->nmodes, tt->nnz, tt->ind, tt->vals, &out, opts) !=
      SPLATT_SUCCESS)
    return NULL;
  return out;
}

/* Coordinate text format: one nonzero per line, N 1-based indices then the
 * value; '#' starts a comment line. */
int adatm_read_tns(char const *fname, uint32_t *nmodes_out, uint64_t *nnz_out,
                   splatt_idx_t ***inds_out, splatt_val_t **vals_out)
{
  FILE *f = fopen(fname, "r");
  if (!f) {
    adatm_set_error("cannot open %s", fname);
    return SPLATT_ERROR_BADINPUT;
  }
  char *line = NULL;
  size_t cap = 0;
  ssize_t len;
  uint32_t N = 0;
  uint64_t n = 0, ncap = 1 << 16;
  splatt_idx_t **ind = NULL;
  splatt_val_t *val = NULL;
  int rc = SPLATT_SUCCESS;
  while ((len = getline(&line, &cap, f)) >= 0) {
    char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '\n' || *p == '\0') continue;
    double tok[MAXN + 2];
    uint32_t nt = 0;
    char *end;
    while (nt < MAXN + 2) {
      double const v = strtod(p, &end);
      if (end == p) break;
      tok[nt++] = v;
      p = end;
    }
    if (N == 0) {
      if (nt < 4 || nt > MAXN + 1) {
        adatm_set_error("%s: %u columns per line", fname, nt);
        rc = SPLATT_ERROR_BADINPUT;
        break;
      }
      N = nt - 1;
      ind = (splatt_idx_t **)calloc(N, sizeof(splatt_idx_t *));
      for (uint32_t m = 0; m < N; ++m)
        ind[m] = (splatt_idx_t *)malloc(ncap * sizeof(splatt_idx_t));
      val = (splatt_val_t *)malloc(ncap * sizeof(splatt_val_t));
    }
    if (nt != N + 1) {
      adatm_set_error("%s: line %llu has %u columns, expected %u", fname,
                      (unsigned long long)n + 1, nt, N + 1);
      rc = SPLATT_ERROR_BADINPUT;
      break;
    }
    if (n == ncap) {
      ncap *= 2;
      for (uint32_t m = 0; m < N; ++m)
        ind[m] = (splatt_idx_t *)realloc(ind[m], ncap * sizeof(splatt_idx_t));
      val = (splatt_val_t *)realloc(val, ncap * sizeof(splatt_val_t));
    }
    for (uint32_t m = 0; m < N; ++m) {
      if (tok[m] < 1.0) {
        adatm_set_error("%s: index below 1 on line %llu", fname,
                        (unsigned long long)n + 1);
        rc = SPLATT_ERROR_BADINPUT;
      }
      ind[m][n] = (splatt_idx_t)tok[m] - 1;
    }
    val[n] = tok[N];
    ++n;
    if (rc != SPLATT_SUCCESS) break;
  }
  free(line);
  fclose(f);
  if (rc == SPLATT_SUCCESS && n == 0) {
    adatm_set_error("%s: no nonzeros", fname);
    rc = SPLATT_ERROR_BADINPUT;
  }
  if (rc != SPLATT_SUCCESS) {
    if (ind)
      for (uint32_t m = 0; m < N; ++m) free(ind[m]);
    free(ind);
    free(val);
    return rc;
  }
  *nmodes_out = N;
  *nnz_out = n;
  *inds_out = ind;
  *vals_out = val;
  return SPLATT_SUCCESS;
}

int splatt_csf_load(char const *const fname, splatt_idx_t *nmodes,
                    splatt_csf **tensors, double const *const options)
{
  uint32_t N;
  uint64_t nnz;
  splatt_idx_t **ind;
  splatt_val_t *val;
  int rc = adatm_read_tns(fname, &N, &nnz, &ind, &val);
  if (rc != SPLATT_SUCCESS) return rc;
  rc = splatt_csf_convert(N, nnz, ind, val, tensors, options);
  for (uint32_t m = 0; m < N; ++m) free(ind[m]);
  free(ind);
  free(val);
  if (rc == SPLATT_SUCCESS && nmodes) *nmodes = N;
  return rc;
}

void splatt_free_csf(splatt_csf *tensors, double const *const options)
{
  if (!tensors) return;
  reg_ent *r = reg_remove(tensors);
  uint32_t n;
  if (r) {
    n = r->n_csf;
    adatm_engine_destroy(r->eng);
    adatm_coo_free(&r->coo);
    free(r);
  } else {
    int const alloc = options ? (int)options[SPLATT_OPTION_CSF_ALLOC]
                              : SPLATT_CSF_ONEMODE;
    n = (alloc == SPLATT_CSF_ALLMODE) ? (uint32_t)tensors->nmodes
        : (alloc == SPLATT_CSF_TWOMODE) ? 2 : 1;
  }
  free_csf_array(tensors, n);
}

void splatt_free_kruskal(splatt_kruskal *factored)
{
  if (!factored) return;
  for (splatt_idx_t m = 0; m < factored->nmodes && m < SPLATT_MAX_NMODES; ++m) {
    free(factored->factors[m]);
    factored->factors[m] = NULL;
  }
  free(factored->lambda);
  factored->lambda = NULL;
}

/* ---- engine management ------------------------------------------------------- */
static uint64_t hash_matrix(double const *A, uint64_t n)
{
  uint64_t h = 0x9E3779B97F4A7C15ull ^ n;
  uint64_t const *w = (uint64_t const *)A;
  #pragma omp parallel for reduction(^ : h) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t x = w[i] + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    h ^= x;
  }
  return h;
}

static int ensure_engine(reg_ent *r, uint64_t rank, int prec)
{
  if (r->eng && r->eng_rank == rank && r->eng_prec == prec) return SPLATT_SUCCESS;
  if (r->eng) {
    adatm_engine_destroy(r->eng);
    r->eng = NULL;
  }
  memset(r->fac_on_dev, 0, sizeof(r->fac_on_dev));
  memset(r->memo_ok, 0, sizeof(r->memo_ok));
  if (!r->coo.nnz) {
    adatm_set_error("CSF set built with SPLATT_OPTION_ADATM_DEVICE < 0 (host arrays "
                    "only): MTTKRP and CP-ALS need a GPU");
    return SPLATT_ERROR_BADINPUT;
  }
  /* built on the device from the coordinates */
  int const rc = adatm_engine_create_use(&r->coo, rank, &r->plan, prec, r->device,
                                         r->use_grp, r->use_tag, &r->eng);
  if (rc == SPLATT_SUCCESS) {
    r->eng_rank = (uint32_t)rank;
    r->eng_prec = prec;
  }
  return rc;
}

static int opt_prec(double const *opts)
{
  int const p = opts ? (int)opts[SPLATT_OPTION_ADATM_PRECISION] : SPLATT_ADATM_FP64;
  return (p >= SPLATT_ADATM_FP64 && p <= SPLATT_ADATM_BF16ROWS) ? p
                                                               : SPLATT_ADATM_FP64;
}

/* upload the factors whose contents changed; returns their hashes */
static int sync_factors(reg_ent *r, splatt_val_t **mats, uint32_t skip,
                        uint64_t *hashes)
{
  for (uint32_t m = 0; m < r->N; ++m) {
    if (m == skip || !mats[m]) continue;
    uint64_t const h = hash_matrix(mats[m], r->dims[m] * r->eng_rank);
    hashes[m] = h;
    if (!r->fac_on_dev[m] || r->fac_hash[m] != h) {
      int const rc = adatm_engine_set_factor(r->eng, m, mats[m]);
      if (rc != SPLATT_SUCCESS) return rc;
      r->fac_hash[m] = h;
      r->fac_on_dev[m] = 1;
    }
  }
  return SPLATT_SUCCESS;
}

/* level-d reuse reads the memo at level L = b - min(n_imten, b - d)
 * (src/mttkrp.c:1578-1591); it is current when the factors of levels
 * L+1 .. N-1 are the ones its group's last root pass contracted */
static int memo_current(reg_ent const *r, uint32_t g, uint32_t d,
                        uint64_t const *hashes)
{
  adatm_group const *G = r->plan.grp + g;
  if (!r->memo_ok[g]) return 0;
  uint32_t const bd = G->begin_imten - d;
  uint32_t const L = G->begin_imten - (G->n_imten < bd ? G->n_imten : bd);
  for (uint32_t l = L + 1; l < r->N; ++l)
    if (r->memo_fac[g][G->o_opt[l]] != hashes[G->o_opt[l]]) return 0;
  return 1;
}

static int mttkrp_registered(reg_ent *r, uint32_t mode, splatt_val_t **mats,
                             splatt_val_t *out)
{
  uint64_t hashes[MAXN] = {0};
  uint32_t const g = r->use_grp[mode];
  uint32_t const root = r->plan.grp[g].memo_mode;
  int rc = sync_factors(r, mats, mode, hashes);
  if (rc != SPLATT_SUCCESS) return rc;
  if (r->use_tag[mode] == 1) {
    uint32_t d = 0;
    while (r->plan.grp[g].o_opt[d] != mode) ++d;
    if (!memo_current(r, g, d, hashes)) {
      /* rebuild the memo: a root pass reads every factor but the root's */
      if (!mats[mode]) {
        adatm_set_error("splatt_mttkrp: memo of group %u is stale and the "
                        "factor of mode %u is needed to rebuild it", g, mode);
        return SPLATT_ERROR_BADINPUT;
      }
      rc = sync_factors(r, mats, root, hashes);
      if (rc != SPLATT_SUCCESS) return rc;
      rc = adatm_engine_mttkrp(r->eng, root, NULL);
      if (rc != SPLATT_SUCCESS) return rc;
      memcpy(r->memo_fac[g], hashes, sizeof(hashes));
      r->memo_ok[g] = 1;
    }
  }
  rc = adatm_engine_mttkrp(r->eng, mode, out);
  if (rc == SPLATT_SUCCESS && r->use_tag[mode] == 0 && r->plan.grp[g].n_imten) {
    memcpy(r->memo_fac[g], hashes, sizeof(hashes));
    r->memo_ok[g] = 1;
  }
  return rc;
}

int splatt_mttkrp(splatt_idx_t const mode, splatt_idx_t const ncolumns,
                  splatt_csf const *const tensors, splatt_val_t **matrices,
                  splatt_val_t *const matout, double const *const options)
{
  reg_ent *r = reg_find(tensors);
  if (!r) {
    adatm_set_error("splatt_mttkrp: CSF not created by this library");
    return SPLATT_ERROR_BADINPUT;
  }
  if (mode >= r->N || ncolumns < 1 || !matrices || !matout)
    return SPLATT_ERROR_BADINPUT;
  for (uint32_t m = 0; m < r->N; ++m)
    if (m != mode && !matrices[m]) return SPLATT_ERROR_BADINPUT;
  int const prec = opt_prec(options);
  int rc = ensure_engine(r, ncolumns, prec);
  if (rc != SPLATT_SUCCESS) return rc;
  return mttkrp_registered(r, (uint32_t)mode, matrices, matout);
}

/* ---- AdaTM compatibility entry points ------------------------------------ */
splatt_csf *csf_alloc_adaptive(sptensor_t *const tt,
                               group_properties *const grp_prop,
                               splatt_idx_t const n_grp,
                               double const *const opts, splatt_idx_t *n_csf)
{
  if (!tt || !grp_prop || n_grp < 1 || tt->nmodes > MAXN) return NULL;
  uint32_t const N = (uint32_t)tt->nmodes;
  adatm_plan p;
  memset(&p, 0, sizeof(p));
  p.n_grp = (uint32_t)n_grp;
  for (uint32_t g = 0; g < n_grp; ++g) {
    p.grp[g].memo_mode = (uint32_t)grp_prop[g].memo_mode;
    p.grp[g].begin_imten = (uint32_t)grp_prop[g].begin_imten;
    p.grp[g].n_imten = (uint32_t)grp_prop[g].n_imten;
    for (uint32_t m = 0; m < N; ++m) p.grp[g].o_opt[m] = (uint32_t)grp_prop[g].o_opt[m];
  }
  uint32_t ug[MAXN], ut[MAXN];
  adatm_plan q = p;
  if (opts && (int)opts[SPLATT_OPTION_CSF_ALLOC] == SPLATT_ADAPTIVE_SPACE_EFFICIENT)
    adatm_plan_space_efficient(&p, N, &q, ug, ut); /* memoising groups' CSFs only */
  else
    adatm_plan_use(&p, N, ug, ut);
  adatm_coo t;
  if (adatm_coo_from_sptensor(tt, &t) != SPLATT_SUCCESS) return NULL;
  double *dflt = opts ? NULL : splatt_default_opts();
  splatt_csf *out = NULL;
  uint32_t nc = 0;
  int const rc = build_from_coo(&t, opts ? opts : dflt, &q, &p, ug, ut, &out, &nc);
  adatm_coo_free(&t);
  free(dflt);
  if (rc != SPLATT_SUCCESS) return NULL;
  if (n_csf) *n_csf = nc;
  return out;
}

rcsf_seq_adaptive *rcsf_alloc_adaptive(splatt_csf *const ct,
                                       splatt_idx_t const nfactors,
                                       group_properties *const grp_prop,
                                       splatt_idx_t const n_grp,
                                       splatt_idx_t const n_csf,
                                       double const *const opts,
                                       splatt_idx_t *n_rcsf)
{
  reg_ent *r = reg_find(ct);
  if (!r) return NULL;
  if (ensure_engine(r, nfactors, opt_prec(opts)) != SPLATT_SUCCESS) return NULL;
  rcsf_seq_adaptive *ret =
      (rcsf_seq_adaptive *)calloc(n_csf ? n_csf : 1, sizeof(rcsf_seq_adaptive));
  splatt_idx_t j = 0;
  for (splatt_idx_t g = 0; g < n_grp; ++g) {
    if (grp_prop[g].n_imten == 0) continue;
    uint32_t const b = (uint32_t)grp_prop[g].begin_imten;
    uint32_t const nim = (uint32_t)grp_prop[g].n_imten;
    ret[j].begin_imten = b;
    ret[j].n_imten = nim;
    ret[j].rcsfs = (splatt_csf *)calloc(nim, sizeof(splatt_csf));
    /* TIME_EFFICIENT: CSF g; SPACE_EFFICIENT: the j-th memoising group's */
    splatt_csf const *c = ct + ((n_csf == n_grp) ? g : j);
    for (uint32_t k = 0; k < nim; ++k) {
      /* vCSF k describes the memo of level b-1-k (adatm_rcsf.c:227-276);
         structure aliases the CSF, values live on the device */
      splatt_csf *v = ret[j].rcsfs + k;
      uint32_t const lvl = b - 1 - k;
      uint32_t const nm = lvl + 2;
      v->nmodes = nm;
      for (uint32_t m = 0; m < nm; ++m) {
        v->dims[m] = c->dims[c->dim_perm[m]];
        v->dim_perm[m] = c->dim_perm[m];
        v->tile_dims[m] = 1;
      }
      v->dims[nm - 1] = nfactors;
      v->which_tile = SPLATT_NOTILE;
      v->ntiles = 1;
      v->nnz = c->pt->nfibs[lvl] * nfactors;
      v->pt = (csf_sparsity *)calloc(1, sizeof(csf_sparsity));
      for (uint32_t m = 0; m + 1 < nm; ++m) v->pt->nfibs[m] = c->pt->nfibs[m];
      v->pt->nfibs[nm - 1] = v->nnz;
      for (uint32_t m = 0; m + 2 < nm; ++m) {
        v->pt->fptr[m] = c->pt->fptr[m];
        v->pt->fids[m] = c->pt->fids[m];
      }
      v->pt->fids[nm - 2] = c->pt->fids[nm - 2];
      v->pt->vals = NULL;
    }
    ++j;
  }
  if (n_rcsf) *n_rcsf = j;
  return ret;
}

size_t csf_storage_adaptive(splatt_csf const *const tensors,
                            splatt_idx_t const n_csf,
                            group_properties *const grp_prop, int const n_grp)
{
  (void)grp_prop;
  (void)n_grp;
  size_t bytes = 0;
  for (splatt_idx_t g = 0; g < n_csf; ++g) {
    splatt_csf const *c = tensors + g;
    bytes += c->nnz * sizeof(splatt_val_t) + c->nnz * sizeof(splatt_idx_t) +
             c->ntiles * sizeof(csf_sparsity);
    for (splatt_idx_t m = 0; m + 1 < c->nmodes; ++m) {
      bytes += (c->pt->nfibs[m] + 1) * sizeof(splatt_idx_t);
      if (c->pt->fids[m]) bytes += c->pt->nfibs[m] * sizeof(splatt_idx_t);
    }
  }
  return bytes;
}

size_t rcsf_storage_adaptive(rcsf_seq_adaptive const *const seq_rcsf,
                             splatt_idx_t const n_rcsf)
{
  size_t bytes = 0;
  for (splatt_idx_t s = 0; s < n_rcsf; ++s)
    for (splatt_idx_t k = 0; k < seq_rcsf[s].n_imten; ++k)
      bytes += seq_rcsf[s].rcsfs[k].nnz * sizeof(splatt_val_t) +
               seq_rcsf[s].rcsfs[k].ntiles * sizeof(csf_sparsity);
  return bytes;
}

void mttkrp_csf_adaptive(splatt_csf const *const tensors,
                         rcsf_seq_adaptive const *const rs_seq,
                         splatt_idx_t const n_csf, matrix_t **mats,
                         splatt_idx_t const mode, void *const thds,
                         group_properties *const grp_prop, int const n_grp,
                         splatt_idx_t const use_csf, splatt_idx_t const use_tag,
                         double const *const opts)
{
  (void)rs_seq; (void)n_csf; (void)thds; (void)grp_prop; (void)n_grp;
  (void)use_csf; (void)use_tag;
  reg_ent *r = reg_find(tensors);
  if (!r || !mats || mode >= r->N) {
    fprintf(stderr, "mttkrp_csf_adaptive: unknown CSF or bad mode\n");
    return;
  }
  splatt_val_t *m[MAXN];
  for (uint32_t k = 0; k < r->N; ++k) m[k] = mats[k] ? mats[k]->vals : NULL;
  matrix_t *M = mats[SPLATT_MAX_NMODES];
  M->I = tensors[0].dims[mode];
  int const rc = splatt_mttkrp(mode, mats[0]->J, tensors, m, M->vals, opts);
  if (rc != SPLATT_SUCCESS)
    fprintf(stderr, "mttkrp_csf_adaptive: %s\n", adatm_last_error());
}

/* ---- CP-ALS ------------------------------------------------------------------ */
static int plan_matches(reg_ent const *r, group_properties const *grp_prop,
                        splatt_idx_t n_grp, splatt_idx_t const *use_csfs,
                        splatt_idx_t const *use_tags);

/* columns to unit 2-norm with the norms folded into lambda (SPLATT's
 * cpd_post_process) */
static void post_process(uint32_t N, uint64_t const *dims, uint64_t R,
                         double **A, double *lambda)
{
  for (uint32_t m = 0; m < N; ++m)
    for (uint64_t r = 0; r < R; ++r) {
      double s = 0.0;
      for (uint64_t i = 0; i < dims[m]; ++i) s += A[m][i * R + r] * A[m][i * R + r];
      s = sqrt(s);
      if (s > 0.0)
        for (uint64_t i = 0; i < dims[m]; ++i) A[m][i * R + r] /= s;
      lambda[r] *= s;
    }
}

/* CP-ALS on a registered CSF set (cpd_als_iterate_adaptive,
 * src/adatm_cpd.c:256-374): initial factors from `init` (caller matrices,
 * dims[m] x rank) or, when NULL, from the seeded generator; the final
 * factors (unit 2-norm columns) go to out[m], their scales to lambda */
static int cpd_core(reg_ent *r, splatt_idx_t nfactors, double const *opts,
                    double *const *init, double **out, double *lambda, double *fit_out)
{
  int const prec = opt_prec(opts);
  int rc = ensure_engine(r, nfactors, prec);
  if (rc != SPLATT_SUCCESS) return rc;
  if (init) {
    for (uint32_t m = 0; rc == SPLATT_SUCCESS && m < r->N; ++m)
      rc = adatm_engine_set_factor(r->eng, m, init[m]);
  } else {
    rc = adatm_engine_init_factors(r->eng, (uint64_t)opts[SPLATT_OPTION_RANDSEED]);
  }
  if (rc != SPLATT_SUCCESS) return rc;
  memset(r->fac_on_dev, 0, sizeof(r->fac_on_dev));
  memset(r->memo_ok, 0, sizeof(r->memo_ok));
  uint32_t const niters = (uint32_t)opts[SPLATT_OPTION_NITER];
  double const tol = opts[SPLATT_OPTION_TOLERANCE];
  int const graph = (int)opts[SPLATT_OPTION_ADATM_GRAPH];
  int const verb = (int)opts[SPLATT_OPTION_VERBOSITY];
  double fit = 0.0, oldfit = 0.0;
  rc = adatm_engine_cpd_begin(r->eng);
  if (rc != SPLATT_SUCCESS) return rc;
  double const t_all = adatm_wtime();
  for (uint32_t it = 0; it < niters; ++it) {
    double const t0 = adatm_wtime();
    rc = adatm_engine_cpd_sweep(r->eng, it, graph, &fit);
    if (rc != SPLATT_SUCCESS) return rc;
    if (verb > SPLATT_VERBOSITY_NONE)
      printf("  its = %3u (%0.3fs)  fit = %0.5f  delta = %+0.4e\n", it + 1,
             adatm_wtime() - t0, fit, fit - oldfit);
    if (it > 0 && fabs(fit - oldfit) < tol) break;
    oldfit = fit;
  }
  if (verb > SPLATT_VERBOSITY_NONE) printf("CPD = %0.3fs\n", adatm_wtime() - t_all);
  int st = 0;
  if (adatm_engine_solve_status(r->eng, &st) == SPLATT_SUCCESS && st == 1 &&
      verb > SPLATT_VERBOSITY_NONE)
    fprintf(stderr, "warning: the CP-ALS normal equations were not numerically positive "
                    "definite; solved with a ridge\n");
  rc = adatm_engine_get_lambda(r->eng, lambda);
  for (uint32_t m = 0; rc == SPLATT_SUCCESS && m < r->N; ++m)
    rc = adatm_engine_get_factor(r->eng, m, out[m]);
  if (rc != SPLATT_SUCCESS) return rc;
  post_process(r->N, r->dims, nfactors, out, lambda);
  *fit_out = fit;
  return SPLATT_SUCCESS;
}

static int cpd_registered(reg_ent *r, splatt_idx_t nfactors,
                          double const *opts, splatt_kruskal *out)
{
  memset(out, 0, sizeof(*out));
  out->rank = nfactors;
  out->nmodes = r->N;
  out->lambda = (splatt_val_t *)malloc(nfactors * sizeof(splatt_val_t));
  if (!out->lambda) return SPLATT_ERROR_NOMEMORY;
  for (uint32_t m = 0; m < r->N; ++m) {
    out->dims[m] = r->dims[m];
    out->factors[m] = (splatt_val_t *)malloc(r->dims[m] * nfactors * sizeof(splatt_val_t));
    if (!out->factors[m]) {
      splatt_free_kruskal(out);
      return SPLATT_ERROR_NOMEMORY;
    }
  }
  double fit = 0.0;
  int const rc = cpd_core(r, nfactors, opts, NULL, out->factors, out->lambda, &fit);
  if (rc != SPLATT_SUCCESS) {
    splatt_free_kruskal(out);
    return rc;
  }
  out->fit = fit;
  return SPLATT_SUCCESS;
}

int splatt_cpd_als(splatt_csf const *const tensors, splatt_idx_t const nfactors,
                   double const *const options, splatt_kruskal *factored)
{
  reg_ent *r = reg_find(tensors);
  if (!r || !factored || nfactors < 1) {
    adatm_set_error("splatt_cpd_als: CSF not created by this library");
    return SPLATT_ERROR_BADINPUT;
  }
  double *dflt = options ? NULL : splatt_default_opts();
  int const rc = cpd_registered(r, nfactors, options ? options : dflt, factored);
  free(dflt);
  return rc;
}

int splatt_cpd_als_adaptive(splatt_csf const *const tensors,
                            rcsf_seq_adaptive *const rs_seq,
                            splatt_idx_t const n_csf,
                            group_properties const *const grp_prop,
                            splatt_idx_t const n_grp,
                            splatt_idx_t const *const use_csfs,
                            splatt_idx_t const *const use_tags,
                            splatt_idx_t const nfactors,
                            double const *const options,
                            splatt_kruskal *factored)
{
  (void)rs_seq; (void)n_csf;
  reg_ent *r = reg_find(tensors);
  if (!r) {
    adatm_set_error("splatt_cpd_als_adaptive: CSF not created by this library");
    return SPLATT_ERROR_BADINPUT;
  }
  if (!plan_matches(r, grp_prop, n_grp, use_csfs, use_tags)) {
    adatm_set_error("splatt_cpd_als_adaptive: the plan differs from the CSFs'");
    return SPLATT_ERROR_BADINPUT;
  }
  return splatt_cpd_als(tensors, nfactors, options, factored);
}

/* the registered plan is the one the arguments describe */
static int plan_matches(reg_ent const *r, group_properties const *grp_prop,
                        splatt_idx_t n_grp, splatt_idx_t const *use_csfs,
                        splatt_idx_t const *use_tags)
{
  if (!grp_prop || !use_csfs || !use_tags || n_grp != r->full.n_grp) return 0;
  for (splatt_idx_t g = 0; g < n_grp; ++g)
    if (grp_prop[g].memo_mode != r->full.grp[g].memo_mode ||
        grp_prop[g].n_imten != r->full.grp[g].n_imten)
      return 0;
  for (uint32_t m = 0; m < r->N; ++m)
    if (use_csfs[m] != r->use_grp[m] || use_tags[m] != r->use_tag[m]) return 0;
  return 1;
}

/* The ALS loop of the reference (adatm_cpd.h:29-41, src/adatm_cpd.c:256-374)
 * on caller matrices: mats[m] (dims[m] x nfactors, row-major) hold the
 * initial factors and receive the final ones (unit 2-norm columns, the
 * norms in lambda); mats[SPLATT_MAX_NMODES] is the reference's MTTKRP
 * buffer and is not used.  `rinfo` (SPLATT's MPI rank information) is
 * ignored and may be NULL.  Returns the final fit, or NaN on failure
 * (message in adatm_last_error()). */
double cpd_als_iterate_adaptive(splatt_csf const *const tensors,
                                rcsf_seq_adaptive *const rs_seq,
                                splatt_idx_t const n_csf,
                                group_properties const *const grp_prop,
                                splatt_idx_t const n_grp,
                                splatt_idx_t const *const use_csfs,
                                splatt_idx_t const *const use_tags, matrix_t **mats,
                                splatt_val_t *const lambda, splatt_idx_t const nfactors,
                                rank_info *const rinfo, double const *const opts)
{
  (void)rs_seq; (void)n_csf; (void)rinfo;
  reg_ent *r = reg_find(tensors);
  if (!r || !mats || !lambda || nfactors < 1) {
    adatm_set_error("cpd_als_iterate_adaptive: CSF not created by this library");
    return NAN;
  }
  if (!plan_matches(r, grp_prop, n_grp, use_csfs, use_tags)) {
    adatm_set_error("cpd_als_iterate_adaptive: the plan differs from the CSFs'");
    return NAN;
  }
  double *A[MAXN];
  for (uint32_t m = 0; m < r->N; ++m) {
    if (!mats[m] || mats[m]->I != r->dims[m] || mats[m]->J != nfactors) {
      adatm_set_error("cpd_als_iterate_adaptive: factor %u is not %llu x %llu", m,
                      (unsigned long long)r->dims[m], (unsigned long long)nfactors);
      return NAN;
    }
    A[m] = mats[m]->vals;
  }
  double *dflt = opts ? NULL : splatt_default_opts();
  double fit = 0.0;
  int const rc = cpd_core(r, nfactors, opts ? opts : dflt, A, A, lambda, &fit);
  free(dflt);
  return rc == SPLATT_SUCCESS ? fit : NAN;
}

/* ---- SPLATT-internal helpers the AdaTM drivers call ---------------------- */
/* tt_read (cmd_bench.c:234, cmd_cpd.c:139): a coordinate text file as read by
 * splatt_csf_load; dims = largest index + 1 per mode */
sptensor_t *tt_read(char const *const fname)
{
  uint32_t N;
  uint64_t nnz;
  splatt_idx_t **ind;
  splatt_val_t *val;
  if (adatm_read_tns(fname, &N, &nnz, &ind, &val) != SPLATT_SUCCESS) return NULL;
  sptensor_t *tt = (sptensor_t *)calloc(1, sizeof(sptensor_t));
  splatt_idx_t *dims = (splatt_idx_t *)calloc(N, sizeof(splatt_idx_t));
  if (!tt || !dims) {
    for (uint32_t m = 0; m < N; ++m) free(ind[m]);
    free(ind);
    free(val);
    free(tt);
    free(dims);
    return NULL;
  }
  for (uint32_t m = 0; m < N; ++m)
    for (uint64_t j = 0; j < nnz; ++j)
      if (ind[m][j] + 1 > dims[m]) dims[m] = ind[m][j] + 1;
  tt->nnz = nnz;
  tt->nmodes = N;
  tt->dims = dims;
  tt->ind = ind;
  tt->vals = val;
  return tt;
}

void tt_free(sptensor_t *tt)
{
  if (!tt) return;
  for (splatt_idx_t m = 0; m < tt->nmodes; ++m) free(tt->ind[m]);
  free(tt->ind);
  free(tt->vals);
  free(tt->dims);
  free(tt);
}

/* csf_free (adatm_bench.c:650) and splatt_csf_free (cmd_cpd.c:255) are
 * SPLATT's internal names of splatt_free_csf */
void csf_free(splatt_csf *tensors, double const *const opts)
{
  splatt_free_csf(tensors, opts);
}

void splatt_csf_free(splatt_csf *tensors, double const *const opts)
{
  splatt_free_csf(tensors, opts);
}

/* zero-filled row-major I x J matrix (SPLATT matrix.h mat_alloc) */
matrix_t *mat_alloc(splatt_idx_t const nrows, splatt_idx_t const ncols)
{
  matrix_t *M = (matrix_t *)malloc(sizeof(matrix_t));
  if (!M) return NULL;
  M->I = nrows;
  M->J = ncols;
  M->rowmajor = 1;
  size_t const n = (size_t)(nrows * ncols);
  M->vals = (splatt_val_t *)calloc(n > 0 ? n : 1, sizeof(splatt_val_t));
  if (!M->vals) {
    free(M);
    return NULL;
  }
  return M;
}

void mat_free(matrix_t *M)
{
  if (!M) return;
  free(M->vals);
  free(M);
}

/* The engine behind a registered CSF (NULL before the rank is known). */
adatm_engine *adatm_csf_engine(splatt_csf const *tensors)
{
  reg_ent *r = reg_find(tensors);
  return r ? r->eng : NULL;
}
