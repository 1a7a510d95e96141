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
#ifndef RED_TILE
#define RED_TILE 128u /* children per reduction work item */
#endif
#ifndef SCAT_TILE
#define SCAT_TILE 64u /* nodes per scatter work item */
#endif

typedef struct {
  adatm_kwork kw; /* device pointers */
  uint32_t nslots;
} dwork;

/* one fused reduction pass of a group: level lb (children) -> lt (output);
 * a leaf pass may also write memo levels between them (side levels) */
typedef struct {
  uint32_t lb, lt;
  dwork w;
  uint32_t side[ADATM_KMAXFUSE + 1]; /* relative level j is written beside the top */
  dwork sw[ADATM_KMAXFUSE + 1];      /* merge of the partial rows of its cut nodes */
  size_t soff[ADATM_KMAXFUSE + 1];   /* its first partial row in the engine's spart */
  uint32_t *shead, *stail;           /* per item and relative level: partial slots */
} dpass;

/* reuse scatter of level d: id-sorted nodes with their ancestors' rows */
typedef struct {
  uint32_t d;
  uint32_t *node;
  uint32_t *rid[ADATM_KMAXSCAT];
  uint32_t *rowptr;
  dwork w;
} dscat;

#define MAXPASS 32

typedef struct {
  uint32_t memo_mode, b, nim;
  uint32_t order[MAXN];
  uint64_t F[MAXN];
  void *vals;
  uint32_t *ids[MAXN];
  uint32_t *ptr[MAXN];
  uint8_t *cfd[MAXN]; /* first differing level after each node (pass bottoms) */
  void *memo[MAXN];
  dpass pass[MAXPASS];
  uint32_t npass;
  dscat scat[MAXN]; /* by level */
  uint32_t *orow;   /* root node -> output row; NULL when they coincide */
  int zero_root;    /* output rows without a root node: zero M first */
} dgroup;

enum { KC_LEAF = 0, KC_LEVEL = 1, KC_SCATTER = 2, KC_DENSE = 3, KC_MERGE = 4, KC_NCLASS = 5 };

/* launches recorded as stages of a persistent launch instead of launched */
typedef struct {
  adatm_kstage *st;
  double *cost;     /* per stage: algorithmic bytes + a per-block overhead */
  uint32_t n, cap;
  double next_cost; /* the bytes of the launch being recorded */
} krec;

#define REC_BLOCK_COST 4096.0 /* bytes-equivalent overhead of one task */

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
  uint32_t red_tile, scat_tile; /* children / nodes per work item */
  uint32_t batch;               /* children per batch of the reductions (0: default) */
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
  void *carry, *carry2;
  size_t tmp_bytes, carry_bytes, carry2_bytes;
  void *spart;       /* partial rows of side-level nodes cut by work items */
  size_t spart_bytes;
  uint32_t flags;    /* ADATM_ENGINE_STAGED, ADATM_ENGINE_STAGED_GROUP(g) */
  double ttnormsq;
  size_t dev_bytes;
  uint32_t fits_cap;
  uint32_t nsweep; /* sweeps since cpd_prepare: the device counter `iter` */
  krec *rec;       /* non-NULL: the MTTKRP launches are recorded, not made */
  hipGraphExec_t gexec;
  hipGraph_t graph;
  int gvalid;
  /* optional per-class timing of launches */
  int instrument;
  hipEvent_t ev[2048];
  int ev_class[1024];
  int ev_info[1024]; /* levels fused (reduce) or scatter level */
  double ev_bytes[1024];
  int nev;
};

/* ---- helpers ------------------------------------------------------------ */
static int dwork_fixlists(adatm_engine *e, dwork *d);
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

static void free_work(dwork *d)
{
  hipFree((void *)d->kw.anc);
  hipFree((void *)d->kw.start);
  hipFree((void *)d->kw.seg);
  hipFree((void *)d->kw.slot);
  hipFree((void *)d->kw.hseg);
  hipFree((void *)d->kw.hslot);
  hipFree((void *)d->kw.cstart);
  hipFree((void *)d->kw.cseg);
  hipFree((void *)d->kw.hchunk);
  hipFree((void *)d->kw.mlist);
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
static void ev_begin(adatm_engine *e, int cls, int info, double bytes)
{
  if (e->rec) e->rec->next_cost = bytes;
  if (!e->instrument || e->nev >= 1024) return;
  hipEventRecord(e->ev[2 * e->nev], e->stream);
  e->ev_class[e->nev] = cls;
  e->ev_info[e->nev] = info;
  e->ev_bytes[e->nev] = bytes;
}

static void ev_end(adatm_engine *e)
{
  if (!e->instrument || e->nev >= 1024) return;
  hipEventRecord(e->ev[2 * e->nev + 1], e->stream);
  e->nev++;
}

/* Compulsory one-touch bytes of a fused pass lb -> lt: bottom input, the
 * structure of every level it walks, each factor once, the output once
 * (4-byte indices on the device). */
static double bytes_pass(adatm_engine const *e, dgroup const *g, uint32_t lb,
                         uint32_t lt)
{
  double const z = (double)e->vsz, w = (double)e->wsz, R = (double)e->R;
  double b;
  if (lb == e->N - 1) /* leaf values, leaf ids, first-differing levels */
    b = (double)g->F[lb] * (z + 4.0 + 1.0);
  else /* memo rows, their node ids, first-differing levels */
    b = (double)g->F[lb] * (R * z + 4.0 + 1.0);
  for (uint32_t l = lt + 1; l < lb; ++l) /* intermediate node ids */
    b += (double)g->F[l] * 4.0;
  for (uint32_t l = lt + 1; l <= lb; ++l)
    b += (double)e->dims[g->order[l]] * R * w;
  b += (double)g->F[lt] * R * z; /* output rows */
  if (lt == 0 && g->orow) b += (double)g->F[0] * 4.0;
  dpass const *P = NULL;
  for (uint32_t i = 0; i < g->npass; ++i)
    if (g->pass[i].lb == lb && g->pass[i].lt == lt) P = g->pass + i;
  for (uint32_t j = 1; P && j < lb - lt; ++j)
    if (P->side[j]) b += (double)g->F[lt + j] * R * z; /* memo rows written beside */
  return b;
}

static double bytes_scatter(adatm_engine const *e, dgroup const *g, uint32_t d)
{
  double const z = (double)e->vsz, w = (double)e->wsz, R = (double)e->R;
  double const Id = (double)e->dims[g->order[d]];
  double const zrow = (d == e->N - 1) ? z : R * z;
  double b = (double)g->F[d] * (zrow + 4.0 * (d + 1)) + (Id + 1.0) * 4.0 +
             Id * R * z;
  for (uint32_t l = 0; l < d; ++l) b += (double)e->dims[g->order[l]] * R * w;
  return b;
}

/* merge of split parents: every carry row read once, chunk rows of the
 * multi-chunk parents written and read once, one output row per parent */
static double merge_bytes(adatm_engine const *e, dwork const *d)
{
  double const row = (double)e->R * (double)e->vsz;
  return ((double)d->nslots + 2.0 * d->kw.nchunk + d->kw.nheavy) * row;
}

/* ---- launch or record ------------------------------------------------------
 * Stage block counts are those of the persistent launch's layout (one lane
 * per column: 64/R items per wave, four waves per block). */
static adatm_kstage *rec_add(adatm_engine *e, uint32_t type, uint64_t nblocks)
{
  krec *r = e->rec;
  if (nblocks > 0xFFFFFFFFull) {
    adatm_set_error("persistent launch: a stage of %llu blocks",
                    (unsigned long long)nblocks);
    return NULL;
  }
  if (r->n == r->cap) {
    uint32_t const cap = r->cap ? 2 * r->cap : 64;
    adatm_kstage *st = (adatm_kstage *)realloc(r->st, (size_t)cap * sizeof(adatm_kstage));
    if (st) r->st = st;
    double *c = (double *)realloc(r->cost, (size_t)cap * sizeof(double));
    if (c) r->cost = c;
    if (!st || !c) {
      adatm_set_error("persistent launch: out of host memory");
      return NULL;
    }
    r->cap = cap;
  }
  r->cost[r->n] = r->next_cost + REC_BLOCK_COST * (double)nblocks;
  r->next_cost = 0.0;
  adatm_kstage *x = r->st + r->n++;
  memset(x, 0, sizeof(*x));
  x->type = type;
  x->nblocks = (uint32_t)nblocks;
  x->dep = ADATM_KS_NODEP;
  x->R = e->R;
  x->gpw = 64u / e->R;
  return x;
}

static uint64_t rec_blocks(adatm_engine const *e, uint64_t n)
{
  uint64_t const per = 4ull * (64u / e->R);
  return (n + per - 1) / per;
}

static int eng_zero(adatm_engine *e, void *p, size_t nbytes)
{
  if (!e->rec) {
    KCHK(adatm_k_zero(e->stream, p, nbytes), "zero M");
    return SPLATT_SUCCESS;
  }
  uint64_t const nw = nbytes / 4;
  if (nw == 0) return SPLATT_SUCCESS;
  e->rec->next_cost = (double)nbytes;
  adatm_kstage *x = rec_add(e, KS_ZERO, (nw + ADATM_KS_ZERO_WORDS - 1) / ADATM_KS_ZERO_WORDS);
  if (!x) return SPLATT_ERROR_NOMEMORY;
  x->out = p;
  x->nwords = nw;
  return SPLATT_SUCCESS;
}

static int eng_reduce(adatm_engine *e, int leaf, adatm_kwork const *w,
                      adatm_kchain const *ch, void const *in, void *dst)
{
  if (!e->rec) {
    KCHK(adatm_k_reduce(e->stream, e->prec, leaf, w, ch, in, dst, e->carry, e->carry2, e->R),
         "reduce");
    return SPLATT_SUCCESS;
  }
  if (w->nitems == 0) return SPLATT_SUCCESS;
  adatm_kstage *x = rec_add(e, KS_REDUCE, rec_blocks(e, w->nitems));
  if (!x) return SPLATT_ERROR_NOMEMORY;
  x->leaf = (uint32_t)leaf;
  x->K = ch->K;
  x->side = ch->side;
  x->w = *w;
  x->ch = *ch;
  x->in = in;
  x->out = dst;
  x->carry = e->carry;
  x->carry2 = e->carry2;
  return SPLATT_SUCCESS;
}

static int eng_fixup(adatm_engine *e, adatm_kwork const *w, void const *carry, void *out)
{
  if (!e->rec) {
    KCHK(adatm_k_fixup(e->stream, e->prec, w, carry, e->carry2, out, e->R), "merge");
    return SPLATT_SUCCESS;
  }
  if (w->nchunk == 0) return SPLATT_SUCCESS;
  adatm_kstage *x = rec_add(e, KS_FIXCHUNK, rec_blocks(e, w->nchunk));
  if (!x) return SPLATT_ERROR_NOMEMORY;
  x->w = *w;
  x->carry = (void *)carry;
  x->carry2 = e->carry2;
  x->out = out;
  if (w->nmulti == 0) return SPLATT_SUCCESS;
  x = rec_add(e, KS_FIXMULTI, w->nmulti);
  if (!x) return SPLATT_ERROR_NOMEMORY;
  x->w = *w;
  x->carry2 = e->carry2;
  x->out = out;
  return SPLATT_SUCCESS;
}

static int eng_scatter(adatm_engine *e, adatm_kwork const *w, uint32_t const *rowptr,
                       adatm_kscat const *sc, void const *in, int zscalar)
{
  if (!e->rec) {
    KCHK(adatm_k_scatter(e->stream, e->prec, w, rowptr, sc, in, zscalar, e->M, e->carry,
                         e->carry2, e->R),
         "scatter");
    return SPLATT_SUCCESS;
  }
  if (w->nitems == 0) return SPLATT_SUCCESS;
  adatm_kstage *x = rec_add(e, KS_SCATTER, rec_blocks(e, w->nitems));
  if (!x) return SPLATT_ERROR_NOMEMORY;
  x->d = sc->d;
  x->zs = (uint32_t)zscalar;
  x->w = *w;
  x->sc = *sc;
  x->rowptr = rowptr;
  x->in = in;
  x->out = e->M;
  x->carry = e->carry;
  x->carry2 = e->carry2;
  return SPLATT_SUCCESS;
}

/* ---- one MTTKRP ----------------------------------------------------------- */
static dpass *find_pass(dgroup *g, uint32_t lb, uint32_t lt)
{
  for (uint32_t i = 0; i < g->npass; ++i)
    if (g->pass[i].lb == lb && g->pass[i].lt == lt) return g->pass + i;
  return NULL;
}

static int run_pass(adatm_engine *e, dgroup *g, uint32_t lb, uint32_t lt,
                    void const *in, void *dst)
{
  dpass *P = find_pass(g, lb, lt);
  if (!P) {
    adatm_set_error("engine: no work list for pass %u -> %u", lb, lt);
    return SPLATT_ERROR_BADINPUT;
  }
  adatm_kchain ch;
  memset(&ch, 0, sizeof(ch));
  ch.K = lb - lt;
  ch.lt = lt;
  ch.rb = e->batch & 0xffu;
  ch.db = (e->batch & ADATM_BATCH_PIPELINED) ? 1u : 0u;
  ch.wide = (e->batch & ADATM_BATCH_WIDE) ? 1u : 0u;
  ch.orow = (lt == 0) ? g->orow : NULL;
  ch.cfd = g->cfd[lb];
  /* node rows through raw buffer loads when every intermediate factor
     spans less than 2^31 bytes (their offsets then fit in 32 bits) */
  ch.nbuf = (e->flags & ADATM_ENGINE_GLOBAL_NODE_ROWS) ? 0u : 1u;
  for (uint32_t j = 1; j <= ch.K; ++j) {
    ch.ids[j] = g->ids[lt + j];
    ch.U[j] = gather_factor(e, g->order[lt + j]);
    if (j < ch.K &&
        (double)e->dims[g->order[lt + j]] * e->R * e->wsz >= (double)0x7fffff00u)
      ch.nbuf = 0u;
  }
  for (uint32_t j = 1; j < ch.K; ++j)
    if (P->side[j]) {
      ch.side = 1;
      ch.sout[j] = g->memo[lt + j];
      ch.spart[j] = (char *)e->spart + P->soff[j] * e->R * e->vsz;
    }
  ch.shead = P->shead;
  ch.stail = P->stail;
  int const leaf = (lb == e->N - 1);
  ev_begin(e, leaf ? KC_LEAF : KC_LEVEL, (int)ch.K, bytes_pass(e, g, lb, lt));
  int rc = eng_reduce(e, leaf, &P->w.kw, &ch, in, dst);
  if (rc != SPLATT_SUCCESS) return rc;
  ev_end(e);
  if (P->w.kw.nchunk) {
    ev_begin(e, KC_MERGE, 0, merge_bytes(e, &P->w));
    rc = eng_fixup(e, &P->w.kw, e->carry, dst);
    if (rc != SPLATT_SUCCESS) return rc;
    ev_end(e);
  }
  /* side levels: the cut nodes' partial rows summed in item order */
  for (uint32_t j = 1; j < ch.K; ++j)
    if (P->side[j] && P->sw[j].kw.nchunk) {
      ev_begin(e, KC_MERGE, 0, merge_bytes(e, &P->sw[j]));
      rc = eng_fixup(e, &P->sw[j].kw, ch.spart[j], g->memo[lt + j]);
      if (rc != SPLATT_SUCCESS) return rc;
      ev_end(e);
    }
  return SPLATT_SUCCESS;
}

/* top of a root-chain pass from lb: staged, the next materialised level at
 * most ADATM_KMAXFUSE levels up; otherwise a leaf pass goes on through the
 * memo levels (written beside its top) up to ADATM_KMAXFUSE levels */
static uint32_t pass_top(adatm_engine const *e, dgroup const *g, uint32_t lb,
                         uint32_t floor_lvl)
{
  uint32_t const gi = (uint32_t)(g - e->g);
  if (!(e->flags & (ADATM_ENGINE_STAGED | ADATM_ENGINE_STAGED_GROUP(gi))) &&
      lb == e->N - 1)
    return lb - floor_lvl > ADATM_KMAXFUSE ? lb - ADATM_KMAXFUSE : floor_lvl;
  uint32_t lt = lb - 1;
  while (lt > floor_lvl && !is_memo_level(g, lt) && lb - lt < ADATM_KMAXFUSE)
    --lt;
  return lt;
}

static int mttkrp_root(adatm_engine *e, dgroup *g, uint32_t mode)
{
  if (g->zero_root) {
    int const rc = eng_zero(e, e->M, e->dims[mode] * e->R * e->vsz);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  void const *in = g->vals;
  int pp = 0;
  uint32_t lb = e->N - 1;
  while (lb > 0) {
    uint32_t const lt = pass_top(e, g, lb, 0);
    void *dst;
    if (lt == 0)
      dst = e->M;
    else if (is_memo_level(g, lt))
      dst = g->memo[lt];
    else {
      dst = e->tmp[pp];
      pp ^= 1;
    }
    int const rc = run_pass(e, g, lb, lt, in, dst);
    if (rc != SPLATT_SUCCESS) return rc;
    in = dst;
    lb = lt;
  }
  return SPLATT_SUCCESS;
}

static uint32_t reuse_source(dgroup const *g, uint32_t d)
{
  uint32_t const bd = g->b - d;
  uint32_t const indeg = g->nim < bd ? g->nim : bd;
  return g->b - indeg;
}

/* Non-root MTTKRP of level d: Z_d from the memo at level reuse_source
 * (tag 1) or from the leaves (tag 2), then the level-d scatter. */
static int mttkrp_level(adatm_engine *e, dgroup *g, uint32_t d, int tag)
{
  uint32_t const N = e->N;
  uint32_t lb = (tag == 1) ? reuse_source(g, d) : N - 1;
  void const *in = (tag == 1) ? g->memo[lb] : g->vals;
  int pp = 0;
  while (lb > d) {
    uint32_t lt = lb > d + ADATM_KMAXFUSE ? lb - ADATM_KMAXFUSE : d;
    void *dst = e->tmp[pp];
    pp ^= 1;
    int const rc = run_pass(e, g, lb, lt, in, dst);
    if (rc != SPLATT_SUCCESS) return rc;
    in = dst;
    lb = lt;
  }
  dscat const *S = g->scat + d;
  adatm_kscat sc;
  memset(&sc, 0, sizeof(sc));
  sc.d = d;
  sc.wide = (e->batch & ADATM_BATCH_WIDE) ? 1u : 0u;
  sc.node = S->node;
  for (uint32_t l = 0; l < d; ++l) {
    sc.rid[l] = S->rid[l];
    sc.U[l] = gather_factor(e, g->order[l]);
  }
  int const zscalar = (d == N - 1); /* leaf level: Z = the values */
  ev_begin(e, KC_SCATTER, (int)d, bytes_scatter(e, g, d));
  int rc = eng_scatter(e, &S->w.kw, S->rowptr, &sc, in, zscalar);
  if (rc != SPLATT_SUCCESS) return rc;
  ev_end(e);
  if (S->w.kw.nchunk) {
    ev_begin(e, KC_MERGE, 0, merge_bytes(e, &S->w));
    rc = eng_fixup(e, &S->w.kw, e->carry, e->M);
    if (rc != SPLATT_SUCCESS) return rc;
    ev_end(e);
  }
  return SPLATT_SUCCESS;
}

static int mttkrp_dev(adatm_engine *e, uint32_t mode)
{
  dgroup *g = e->g + e->use_grp[mode];
  if (e->use_tag[mode] == 0) return mttkrp_root(e, g, mode);
  return mttkrp_level(e, g, e->level_of[e->use_grp[mode]][mode],
                      (int)e->use_tag[mode]);
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
  if ((rc = dalloc(e, (void **)&e->G, (size_t)N * R * R * sizeof(double))) !=
      SPLATT_SUCCESS)
    return rc;
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

static int engine_finish(adatm_engine *e)
{
  int rc;
  if ((rc = dalloc(e, &e->tmp[0], e->tmp_bytes)) != SPLATT_SUCCESS) return rc;
  if ((rc = dalloc(e, &e->tmp[1], e->tmp_bytes)) != SPLATT_SUCCESS) return rc;
  if ((rc = dalloc(e, &e->carry, e->carry_bytes ? e->carry_bytes : 64)) !=
      SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, &e->carry2, e->carry2_bytes ? e->carry2_bytes : 64)) !=
      SPLATT_SUCCESS)
    return rc;
  if ((rc = dalloc(e, &e->spart, e->spart_bytes ? e->spart_bytes : 64)) != SPLATT_SUCCESS)
    return rc;
  return SPLATT_SUCCESS;
}

static int engine_begin(adatm_engine *e, uint32_t N, uint64_t const *dims,
                        uint64_t nnz, uint64_t rank, adatm_plan const *plan,
                        int precision, int device, uint32_t const *use_grp,
                        uint32_t const *use_tag)
{
  if (N < 3 || N > MAXN || rank < 1 || rank > 64 || plan->n_grp < 1 ||
      plan->n_grp > N) {
    adatm_set_error("engine: unsupported shape (N=%u, R=%llu, groups=%u)", N,
                    (unsigned long long)rank, plan->n_grp);
    return SPLATT_ERROR_BADINPUT;
  }
  if (precision < SPLATT_ADATM_FP64 || precision > SPLATT_ADATM_BF16ROWS)
    return SPLATT_ERROR_BADINPUT;
  for (uint32_t m = 0; m < N; ++m)
    if (dims[m] * rank >= (1ull << 32)) { /* factor rows are addressed in 32 bits */
      adatm_set_error("engine: mode %u has %llu x %llu factor entries (limit 2^32)", m,
                      (unsigned long long)dims[m], (unsigned long long)rank);
      return SPLATT_ERROR_BADINPUT;
    }
  e->device = device;
  e->N = N;
  e->R = (uint32_t)rank;
  if (!e->red_tile) e->red_tile = RED_TILE;
  if (!e->scat_tile) e->scat_tile = SCAT_TILE;
  e->prec = precision;
  e->vsz = (precision == SPLATT_ADATM_FP64) ? 8 : 4;
  e->wsz = (precision == SPLATT_ADATM_FP64) ? 8
           : (precision == SPLATT_ADATM_FP32) ? 4 : 2;
  e->nnz = nnz;
  memcpy(e->dims, dims, N * sizeof(uint64_t));
  e->plan = *plan;
  if (use_grp && use_tag) {
    memcpy(e->use_grp, use_grp, N * sizeof(uint32_t));
    memcpy(e->use_tag, use_tag, N * sizeof(uint32_t));
  } else {
    adatm_plan_use(plan, N, e->use_grp, e->use_tag);
  }
  for (uint32_t gi = 0; gi < plan->n_grp; ++gi)
    for (uint32_t l = 0; l < N; ++l) e->level_of[gi][plan->grp[gi].o_opt[l]] = l;
  /* validate the plan: each reuse level must sit above the group's memo */
  for (uint32_t m = 0; m < N; ++m) {
    uint32_t const gi = e->use_grp[m];
    adatm_group const *pg = plan->grp + gi;
    if (gi >= plan->n_grp || e->use_tag[m] > 2) {
      adatm_set_error("engine: bad use map for mode %u", m);
      return SPLATT_ERROR_BADINPUT;
    }
    if (e->use_tag[m] == 1) {
      uint32_t const d = e->level_of[gi][m];
      if (pg->n_imten == 0 || d < 1 || d >= pg->begin_imten) {
        adatm_set_error("engine: mode %u cannot reuse group %u's memo", m, gi);
        return SPLATT_ERROR_BADINPUT;
      }
    } else if (e->use_tag[m] == 2) {
      if (e->level_of[gi][m] < 1) {
        adatm_set_error("engine: mode %u is group %u's root", m, gi);
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

/* ======================================================================== *
 * Construction on the device: sort, CSF levels, work lists, scatter tuples.
 * ======================================================================== */

/* coordinates resident on the device (indices u32, values FP64) */
typedef struct {
  uint32_t N;
  uint64_t nnz;
  uint64_t dims[MAXN];
  uint32_t *ind[MAXN];
  double *vals;
} dcoo;

static void dcoo_free(dcoo *c)
{
  for (uint32_t m = 0; m < MAXN; ++m) hipFree(c->ind[m]);
  hipFree(c->vals);
  memset(c, 0, sizeof(*c));
}

static int dcoo_upload(dcoo *c, adatm_coo const *t, int with_vals)
{
  memset(c, 0, sizeof(*c));
  c->N = t->nmodes;
  c->nnz = t->nnz;
  memcpy(c->dims, t->dims, t->nmodes * sizeof(uint64_t));
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    HIPCHK(hipMalloc((void **)&c->ind[m], t->nnz * 4), "upload coo");
    HIPCHK(hipMemcpy(c->ind[m], t->ind[m], t->nnz * 4, hipMemcpyHostToDevice),
           "upload coo");
  }
  if (with_vals) {
    HIPCHK(hipMalloc((void **)&c->vals, t->nnz * 8), "upload coo");
    HIPCHK(hipMemcpy(c->vals, t->vals, t->nnz * 8, hipMemcpyHostToDevice),
           "upload coo");
  }
  return SPLATT_SUCCESS;
}

static uint32_t nbits_for(uint64_t dim)
{
  uint32_t b = 0;
  while (b < 32 && (1ull << b) < dim) ++b;
  return b;
}

/* perm := stable lexicographic order of the coordinates by `order` */
static int dcoo_lexsort(hipStream_t s, dcoo const *c, uint32_t const *order,
                        uint32_t *perm, uint32_t *keys, uint32_t *stemp)
{
  uint64_t const n = c->nnz;
  KCHK(adatm_k_iota(s, perm, n), "iota");
  for (int64_t k = (int64_t)c->N - 1; k >= 0; --k) {
    uint32_t const nb = nbits_for(c->dims[order[k]]);
    if (nb == 0) continue;
    KCHK(adatm_k_gather(s, keys, c->ind[order[k]], perm, n), "gather keys");
    KCHK(adatm_k_sort_pairs(s, keys, perm, n, nb, stemp), "radix sort");
  }
  return SPLATT_SUCCESS;
}

static int d2h_u32(hipStream_t s, uint32_t const *dsrc, uint32_t *h)
{
  HIPCHK(hipMemcpyAsync(h, dsrc, 4, hipMemcpyDeviceToHost, s), "download");
  HIPCHK(hipStreamSynchronize(s), "sync");
  return SPLATT_SUCCESS;
}

/* exclusive scan with the total brought back to the host */
static int dscan(hipStream_t s, uint32_t const *in, uint32_t *out, uint64_t n,
                 uint32_t *stemp, uint32_t *tot_dev, uint32_t *total)
{
  KCHK(adatm_k_exscan(s, in, out, n, stemp, tot_dev), "scan");
  return d2h_u32(s, tot_dev, total);
}

/* Fiber counts of the device coordinates under `order` */
static int dcoo_count(hipStream_t s, dcoo const *c, uint32_t const *order,
                      uint64_t *F)
{
  uint64_t const n = c->nnz;
  uint32_t const N = c->N;
  uint32_t *perm = NULL, *keys = NULL, *stemp = NULL, *flags = NULL,
           *pos = NULL, *tot = NULL;
  uint32_t *sc[MAXN] = {0};
  uint8_t *fd = NULL;
  int rc = SPLATT_ERROR_NOMEMORY;
  uint64_t const tw = adatm_k_sort_temp_words(n) + adatm_k_scan_temp_words(n);
  if (hipMalloc((void **)&perm, n * 4) || hipMalloc((void **)&keys, n * 4) ||
      hipMalloc((void **)&stemp, tw * 4) || hipMalloc((void **)&fd, n) ||
      hipMalloc((void **)&flags, n * 4) || hipMalloc((void **)&pos, n * 4) ||
      hipMalloc((void **)&tot, 16))
    goto out;
  for (uint32_t l = 0; l < N; ++l)
    if (hipMalloc((void **)&sc[l], n * 4)) goto out;
  if ((rc = dcoo_lexsort(s, c, order, perm, keys, stemp)) != SPLATT_SUCCESS)
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  for (uint32_t l = 0; l < N; ++l)
    if (adatm_k_gather(s, sc[l], c->ind[order[l]], perm, n)) goto out;
  if (adatm_k_first_diff(s, (uint32_t const *const *)sc, N, n, fd)) goto out;
  for (uint32_t l = 0; l + 1 < N; ++l) {
    uint32_t total = 0;
    if (adatm_k_level_flags(s, fd, n, l, flags)) goto out;
    if (dscan(s, flags, pos, n, stemp, tot, &total) != SPLATT_SUCCESS) goto out;
    F[l] = total;
  }
  F[N - 1] = n;
  rc = SPLATT_SUCCESS;
out:
  hipFree(perm); hipFree(keys); hipFree(stemp); hipFree(fd); hipFree(flags);
  hipFree(pos); hipFree(tot);
  for (uint32_t l = 0; l < N; ++l) hipFree(sc[l]);
  if (rc != SPLATT_SUCCESS && !adatm_last_error()[0])
    adatm_set_error("device fiber count failed");
  return rc;
}

int adatm_count_fibers_device_many(adatm_coo const *t, uint32_t const *orders,
                                   uint32_t norders, uint64_t *F, int device)
{
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
  dcoo c;
  int rc = dcoo_upload(&c, t, 0);
  for (uint32_t k = 0; rc == SPLATT_SUCCESS && k < norders; ++k)
    rc = dcoo_count(s, &c, orders + (size_t)k * t->nmodes, F + (size_t)k * t->nmodes);
  dcoo_free(&c);
  hipStreamDestroy(s);
  return rc;
}

/* One CSF in host memory built on the device (the SPLATT API's arrays):
 * stable radix sort by `order`, first differing levels, per level the node
 * flags, their scan, ids and starts, child pointers; the same kernels and
 * arithmetic as the engine's group build, so the structure is the engine's.
 * Fills h->F, fptr, fids (root ids included) and the values in sorted order;
 * h->par stays NULL. */
int adatm_hcsf_build_device(adatm_coo const *t, uint32_t const *order, int device,
                            adatm_hcsf *h)
{
  memset(h, 0, sizeof(*h));
  uint32_t const N = t->nmodes;
  uint64_t const n = t->nnz;
  if (N < 2 || N > MAXN || n == 0 || n >= (1ull << 32)) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
  dcoo c;
  memset(&c, 0, sizeof(c));
  uint32_t *perm = NULL, *keys = NULL, *stemp = NULL, *tot = NULL, *ptr = NULL;
  uint32_t *sc[MAXN] = {0}, *ids = NULL;
  uint32_t *flags[2] = {0}, *pos[2] = {0}, *start[2] = {0};
  uint8_t *fd = NULL;
  double *v64 = NULL;
  uint32_t Fprev = 0;
  int rc = dcoo_upload(&c, t, 1);
  if (rc != SPLATT_SUCCESS) goto out;
  rc = SPLATT_ERROR_NOMEMORY;
  uint64_t const tw = adatm_k_sort_temp_words(n) + adatm_k_scan_temp_words(n);
  if (hipMalloc((void **)&perm, n * 4) || hipMalloc((void **)&keys, n * 4) ||
      hipMalloc((void **)&stemp, tw * 4) || hipMalloc((void **)&fd, n) ||
      hipMalloc((void **)&tot, 16) || hipMalloc((void **)&v64, n * 8) ||
      hipMalloc((void **)&ids, n * 4) || hipMalloc((void **)&ptr, (n + 1) * 4))
    goto out;
  for (uint32_t l = 0; l < N; ++l)
    if (hipMalloc((void **)&sc[l], n * 4)) goto out;
  for (int k = 0; k < 2; ++k)
    if (hipMalloc((void **)&flags[k], n * 4) || hipMalloc((void **)&pos[k], n * 4) ||
        hipMalloc((void **)&start[k], n * 4))
      goto out;
  h->N = N;
  h->nnz = n;
  memcpy(h->order, order, N * sizeof(uint32_t));
  for (uint32_t l = 0; l < N; ++l) {
    h->fids[l] = (uint32_t *)malloc(n * sizeof(uint32_t)); /* trimmed below */
    if (!h->fids[l]) goto out;
  }
  h->vals = (double *)malloc(n * sizeof(double));
  if (!h->vals) goto out;
  if ((rc = dcoo_lexsort(s, &c, order, perm, keys, stemp)) != SPLATT_SUCCESS) goto out;
  rc = SPLATT_ERROR_DEVICE;
  for (uint32_t l = 0; l < N; ++l)
    if (adatm_k_gather(s, sc[l], c.ind[order[l]], perm, n)) goto out;
  if (adatm_k_gather_vals(s, KP_FP64, v64, c.vals, perm, n)) goto out;
  if (hipMemcpyAsync(h->vals, v64, n * 8, hipMemcpyDeviceToHost, s) != hipSuccess) goto out;
  if (adatm_k_first_diff(s, (uint32_t const *const *)sc, N, n, fd)) goto out;
  for (uint32_t l = 0; l < N; ++l) {
    int const cur = l & 1, prv = cur ^ 1;
    uint32_t Fl = (uint32_t)n;
    if (l + 1 < N) {
      if (adatm_k_level_flags(s, fd, n, l, flags[cur])) goto out;
      if (dscan(s, flags[cur], pos[cur], n, stemp, tot, &Fl) != SPLATT_SUCCESS) goto out;
      if (adatm_k_level_emit(s, flags[cur], pos[cur], sc[l], n, ids, start[cur])) goto out;
      if (hipMemcpyAsync(h->fids[l], ids, (size_t)Fl * 4, hipMemcpyDeviceToHost, s) !=
          hipSuccess)
        goto out;
    } else if (hipMemcpyAsync(h->fids[l], sc[l], n * 4, hipMemcpyDeviceToHost, s) !=
               hipSuccess) {
      goto out;
    }
    h->F[l] = Fl;
    if (l >= 1) {
      /* child pointers of level l-1 into level l */
      if (adatm_k_fptr(s, start[prv], Fprev, l + 1 < N ? pos[cur] : NULL, Fl, ptr)) goto out;
      h->fptr[l - 1] = (uint32_t *)malloc(((size_t)Fprev + 1) * sizeof(uint32_t));
      if (!h->fptr[l - 1] ||
          hipMemcpyAsync(h->fptr[l - 1], ptr, ((size_t)Fprev + 1) * 4,
                         hipMemcpyDeviceToHost, s) != hipSuccess)
        goto out;
    }
    /* the downloads read ids/ptr before the next level overwrites them */
    if (hipStreamSynchronize(s) != hipSuccess) goto out;
    Fprev = Fl;
  }
  for (uint32_t l = 0; l + 1 < N; ++l) {
    uint32_t *shrunk = (uint32_t *)realloc(h->fids[l], (h->F[l] ? h->F[l] : 1) * 4);
    if (shrunk) h->fids[l] = shrunk;
  }
  rc = SPLATT_SUCCESS;
out:
  if (hipStreamSynchronize(s) != hipSuccess && rc == SPLATT_SUCCESS)
    rc = hip_fail(hipGetLastError(), "device CSF build");
  hipFree(perm); hipFree(keys); hipFree(stemp); hipFree(fd); hipFree(tot);
  hipFree(v64); hipFree(ids); hipFree(ptr);
  for (uint32_t l = 0; l < N; ++l) hipFree(sc[l]);
  for (int k = 0; k < 2; ++k) {
    hipFree(flags[k]);
    hipFree(pos[k]);
    hipFree(start[k]);
  }
  dcoo_free(&c);
  hipStreamDestroy(s);
  if (rc != SPLATT_SUCCESS) {
    if (!adatm_last_error()[0]) adatm_set_error("device CSF build failed (%d)", rc);
    adatm_hcsf_free(h);
  }
  return rc;
}

/* Work list of a segmented reduction over segments cp[0..nseg] on the device
 * (light items start where the bucket cp/tile changes, long segments are cut
 * into tile-sized pieces with carry slots). */
static int dwork_build(adatm_engine *e, uint32_t const *cp, uint64_t nseg,
                       uint32_t tile, dwork *d, uint32_t *stemp, uint32_t *tot)
{
  hipStream_t const s = e->stream;
  memset(d, 0, sizeof(*d));
  uint32_t *nitem = NULL, *npiece = NULL, *heavy = NULL, *ibase = NULL,
           *sbase = NULL, *hbase = NULL;
  int rc = SPLATT_ERROR_NOMEMORY;
  uint32_t nitems = 0, nslots = 0, nheavy = 0;
  uint64_t const ns = nseg ? nseg : 1;
  if (hipMalloc((void **)&nitem, ns * 4) || hipMalloc((void **)&npiece, ns * 4) ||
      hipMalloc((void **)&heavy, ns * 4) || hipMalloc((void **)&ibase, ns * 4) ||
      hipMalloc((void **)&sbase, ns * 4) || hipMalloc((void **)&hbase, ns * 4))
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  if (nseg) {
    if (adatm_k_items_count(s, cp, nseg, tile, nitem, npiece, heavy)) goto out;
    if (dscan(s, nitem, ibase, nseg, stemp, tot, &nitems) != SPLATT_SUCCESS) goto out;
    if (dscan(s, npiece, sbase, nseg, stemp, tot, &nslots) != SPLATT_SUCCESS) goto out;
    if (dscan(s, heavy, hbase, nseg, stemp, tot, &nheavy) != SPLATT_SUCCESS) goto out;
  }
  if ((rc = dalloc(e, (void **)&d->kw.start, (nitems + 1) * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&d->kw.seg, (nitems + 1) * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&d->kw.slot, (nitems + 1) * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&d->kw.hslot, (nheavy + 1) * 4)) != SPLATT_SUCCESS)
    goto out;
  if (nheavy &&
      (rc = dalloc(e, (void **)&d->kw.hseg, nheavy * 4)) != SPLATT_SUCCESS)
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  if (nseg &&
      adatm_k_items_emit(s, cp, nseg, tile, nitem, ibase, sbase, heavy, hbase,
                         (uint32_t *)d->kw.start, (uint32_t *)d->kw.seg,
                         (uint32_t *)d->kw.slot, (uint32_t *)d->kw.hseg,
                         (uint32_t *)d->kw.hslot))
    goto out;
  if (adatm_k_items_tail(s, cp, nseg, nitems, nheavy, nslots,
                         (uint32_t *)d->kw.start, (uint32_t *)d->kw.seg,
                         (uint32_t *)d->kw.slot, (uint32_t *)d->kw.hslot))
    goto out;
  d->kw.nitems = nitems;
  d->kw.nheavy = nheavy;
  d->nslots = nslots;
  {
    size_t const cb = (size_t)nslots * e->R * e->vsz;
    if (cb > e->carry_bytes) e->carry_bytes = cb;
  }
  rc = dwork_fixlists(e, d);
out:
  hipStreamSynchronize(s); /* kernels above read the temporaries */
  hipFree(nitem); hipFree(npiece); hipFree(heavy); hipFree(ibase);
  hipFree(sbase); hipFree(hbase);
  return rc;
}

/* Merge lists of the split parents of a work list (hslot on the device):
 * chunks of at most ADATM_FIX_CHUNK carry slots, and the parents that span
 * several chunks. */
static int dwork_fixlists(adatm_engine *e, dwork *d)
{
  hipStream_t const s = e->stream;
  uint32_t const nh = d->kw.nheavy;
  d->kw.nchunk = 0;
  d->kw.nmulti = 0;
  if (nh == 0) return SPLATT_SUCCESS;
  uint32_t *nch = NULL, *multi = NULL, *mbase = NULL, *stemp = NULL, *tot = NULL;
  uint32_t nchunk = 0, nmulti = 0;
  int rc = SPLATT_ERROR_NOMEMORY;
  if (hipMalloc((void **)&nch, (size_t)nh * 4) ||
      hipMalloc((void **)&multi, (size_t)nh * 4) ||
      hipMalloc((void **)&mbase, (size_t)nh * 4) ||
      hipMalloc((void **)&stemp, adatm_k_scan_temp_words(nh) * 4) ||
      hipMalloc((void **)&tot, 16))
    goto out;
  if ((rc = dalloc(e, (void **)&d->kw.hchunk, ((size_t)nh + 1) * 4)) != SPLATT_SUCCESS)
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  if (adatm_k_fix_count(s, d->kw.hslot, nh, ADATM_FIX_CHUNK, nch, multi)) goto out;
  if (dscan(s, nch, (uint32_t *)d->kw.hchunk, nh, stemp, tot, &nchunk) != SPLATT_SUCCESS)
    goto out;
  if (dscan(s, multi, mbase, nh, stemp, tot, &nmulti) != SPLATT_SUCCESS) goto out;
  if ((rc = dalloc(e, (void **)&d->kw.cstart, (size_t)nchunk * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&d->kw.cseg, (size_t)nchunk * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&d->kw.mlist, ((size_t)nmulti + 1) * 4)) != SPLATT_SUCCESS)
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  if (adatm_k_fix_emit(s, d->kw.hslot, nh, ADATM_FIX_CHUNK, (uint32_t *)d->kw.hchunk,
                       nchunk, multi, mbase, (uint32_t *)d->kw.cstart,
                       (uint32_t *)d->kw.cseg, (uint32_t *)d->kw.mlist))
    goto out;
  d->kw.nchunk = nchunk;
  d->kw.nmulti = nmulti;
  {
    size_t const cb = (size_t)nchunk * e->R * e->vsz;
    if (cb > e->carry2_bytes) e->carry2_bytes = cb;
  }
  rc = SPLATT_SUCCESS;
out:
  hipStreamSynchronize(s);
  hipFree(nch); hipFree(multi); hipFree(mbase); hipFree(stemp); hipFree(tot);
  if (rc != SPLATT_SUCCESS && !adatm_last_error()[0])
    adatm_set_error("split-parent merge lists failed");
  return rc;
}

/* construction state of one group on the device */
typedef struct {
  uint32_t *par[MAXN];  /* par[l]: level l -> l-1 */
  uint32_t *root_ids;   /* fids of level 0 */
  uint32_t *stemp, *tot;
} dbuild;

/* Side levels of a leaf pass: for every memo level L it crosses, the nodes
 * its work items' ends cut, their partial-row slots (in item order, so each
 * cut node's slots are consecutive) and the merge list of the cut nodes
 * (node = memo row, slot range), chunked like the split-parent merge. */
static int build_side(adatm_engine *e, dgroup *g, dbuild *B, dpass *P)
{
  hipStream_t const s = e->stream;
  uint32_t const K = P->lb - P->lt, n = P->w.kw.nitems;
  uint32_t *cnt = NULL, *base = NULL, *nf = NULL, *nl = NULL, *fl = NULL;
  uint32_t *snode = NULL, *flags = NULL, *pos = NULL;
  size_t total = 0;
  int rc = SPLATT_ERROR_NOMEMORY;
  if ((rc = dalloc(e, (void **)&P->shead, (size_t)n * K * 4)) != SPLATT_SUCCESS ||
      (rc = dalloc(e, (void **)&P->stail, (size_t)n * K * 4)) != SPLATT_SUCCESS)
    return rc;
  rc = SPLATT_ERROR_NOMEMORY;
  /* cut nodes lie in split top parents only: at most two per piece */
  if (hipMalloc((void **)&cnt, (size_t)n * 4) || hipMalloc((void **)&base, (size_t)n * 4) ||
      hipMalloc((void **)&nf, (size_t)n * 4) || hipMalloc((void **)&nl, (size_t)n * 4) ||
      hipMalloc((void **)&fl, (size_t)n * 4) ||
      hipMalloc((void **)&snode, (size_t)2 * n * 4) ||
      hipMalloc((void **)&flags, (size_t)2 * n * 4) ||
      hipMalloc((void **)&pos, (size_t)2 * n * 4))
    goto out;
  for (uint32_t j = 1; j < K; ++j) {
    if (!P->side[j]) continue;
    uint32_t const L = P->lt + j;
    uint32_t nsl = 0, nrun = 0;
    rc = SPLATT_ERROR_DEVICE;
    if (adatm_k_side_count(s, P->w.kw.start, n, P->lb, L, g->F[P->lb], g->cfd[P->lb],
                           (uint32_t const *const *)B->par, cnt, nf, nl, fl))
      goto out;
    if (dscan(s, cnt, base, n, B->stemp, B->tot, &nsl) != SPLATT_SUCCESS) goto out;
    if (adatm_k_side_emit(s, n, K, j, cnt, base, nf, nl, fl, (uint32_t *)P->shead,
                          (uint32_t *)P->stail, snode))
      goto out;
    P->soff[j] = total;
    total += nsl;
    dwork *W = P->sw + j;
    memset(W, 0, sizeof(*W));
    if (nsl) {
      if (adatm_k_run_flags(s, snode, nsl, flags)) goto out;
      if (dscan(s, flags, pos, nsl, B->stemp, B->tot, &nrun) != SPLATT_SUCCESS) goto out;
      if ((rc = dalloc(e, (void **)&W->kw.hseg, (size_t)nrun * 4)) != SPLATT_SUCCESS ||
          (rc = dalloc(e, (void **)&W->kw.hslot, ((size_t)nrun + 1) * 4)) != SPLATT_SUCCESS)
        goto out;
      rc = SPLATT_ERROR_DEVICE;
      if (adatm_k_run_emit(s, flags, pos, snode, nsl, nrun, (uint32_t *)W->kw.hseg,
                           (uint32_t *)W->kw.hslot))
        goto out;
      W->kw.nheavy = nrun;
      W->nslots = nsl;
      if ((rc = dwork_fixlists(e, W)) != SPLATT_SUCCESS) goto out;
    }
  }
  {
    size_t const sb = total * e->R * e->vsz;
    if (sb > e->spart_bytes) e->spart_bytes = sb;
  }
  rc = SPLATT_SUCCESS;
out:
  hipStreamSynchronize(s);
  hipFree(cnt); hipFree(base); hipFree(nf); hipFree(nl); hipFree(fl);
  hipFree(snode); hipFree(flags); hipFree(pos);
  if (rc != SPLATT_SUCCESS && !adatm_last_error()[0])
    adatm_set_error("side-level cut lists failed");
  return rc;
}

/* work list of the pass lb -> lt; a root-chain leaf pass also writes the
 * memo levels it crosses (side levels) */
static int build_pass_dev(adatm_engine *e, dgroup *g, dbuild *B, uint32_t lb,
                          uint32_t lt, int root)
{
  if (find_pass(g, lb, lt)) return SPLATT_SUCCESS;
  if (g->npass >= MAXPASS) return SPLATT_ERROR_BADINPUT;
  hipStream_t const s = e->stream;
  if (!g->cfd[lb]) {
    int const rc = dalloc(e, (void **)&g->cfd[lb], g->F[lb]);
    if (rc != SPLATT_SUCCESS) return rc;
    KCHK(adatm_k_cfd(s, (uint32_t const *const *)B->par, lb, g->F[lb], g->cfd[lb]),
         "first differing levels");
  }
  uint64_t const np = g->F[lt];
  uint32_t *cp = NULL;
  HIPCHK(hipMalloc((void **)&cp, (np + 1) * 4), "alloc cp");
  HIPCHK(hipMemcpyAsync(cp, g->ptr[lt], (np + 1) * 4, hipMemcpyDeviceToDevice, s),
         "copy cp");
  for (uint32_t l = lt + 1; l < lb; ++l)
    KCHK(adatm_k_compose(s, cp, np + 1, g->ptr[l]), "compose");
  dpass *P = g->pass + g->npass;
  P->lb = lb;
  P->lt = lt;
  int rc = dwork_build(e, cp, np, e->red_tile, &P->w, B->stemp, B->tot);
  hipFree(cp);
  if (rc != SPLATT_SUCCESS) return rc;
  uint32_t const K = lb - lt;
  if (lt == 0 && g->orow && P->w.kw.nheavy) {
    /* split root parents are written through the root ids */
    uint32_t *mapped = NULL;
    HIPCHK(hipMalloc((void **)&mapped, (size_t)P->w.kw.nheavy * 4), "alloc map");
    KCHK(adatm_k_gather(s, mapped, B->root_ids, P->w.kw.hseg, P->w.kw.nheavy),
         "map heavy rows");
    HIPCHK(hipMemcpyAsync((void *)P->w.kw.hseg, mapped, (size_t)P->w.kw.nheavy * 4,
                          hipMemcpyDeviceToDevice, s), "copy map");
    HIPCHK(hipStreamSynchronize(s), "map sync");
    hipFree(mapped);
  }
  if (K > 1 && P->w.kw.nitems) {
    if ((rc = dalloc(e, (void **)&P->w.kw.anc,
                     (size_t)P->w.kw.nitems * (K - 1) * 4)) != SPLATT_SUCCESS)
      return rc;
    KCHK(adatm_k_items_anc(s, P->w.kw.start, P->w.kw.nitems, lb, lt,
                           (uint32_t const *const *)B->par, (uint32_t *)P->w.kw.anc),
         "item ancestors");
  }
  g->npass++; /* from here on the destroy path frees what the pass holds */
  if (root && lb == e->N - 1 && K > 1 && P->w.kw.nitems) {
    int any = 0;
    for (uint32_t j = 1; j < K; ++j) {
      P->side[j] = (uint32_t)is_memo_level(g, lt + j);
      any |= (int)P->side[j];
    }
    if (any) rc = build_side(e, g, B, P);
  }
  return rc;
}

static int build_scatter_dev(adatm_engine *e, dgroup *g, dbuild *B, uint32_t d)
{
  dscat *S = g->scat + d;
  if (S->node) return SPLATT_SUCCESS;
  if (d < 1 || d > ADATM_KMAXSCAT) {
    adatm_set_error("engine: reuse level %u beyond %d", d, ADATM_KMAXSCAT);
    return SPLATT_ERROR_BADINPUT;
  }
  hipStream_t const s = e->stream;
  uint64_t const nrows = e->dims[g->order[d]];
  uint64_t const Fd = g->F[d];
  int rc;
  uint32_t *keys = NULL, *sortmp = NULL;
  HIPCHK(hipMalloc((void **)&keys, Fd * 4), "alloc keys");
  HIPCHK(hipMalloc((void **)&sortmp, adatm_k_sort_temp_words(Fd) * 4), "alloc sort");
  HIPCHK(hipMemcpyAsync(keys, g->ids[d], Fd * 4, hipMemcpyDeviceToDevice, s),
         "copy ids");
  S->d = d;
  if ((rc = dalloc(e, (void **)&S->node, Fd * 4)) != SPLATT_SUCCESS) return rc;
  KCHK(adatm_k_iota(s, S->node, Fd), "iota");
  KCHK(adatm_k_sort_pairs(s, keys, S->node, Fd, nbits_for(nrows), sortmp),
       "sort by row");
  if ((rc = dalloc(e, (void **)&S->rowptr, (nrows + 1) * 4)) != SPLATT_SUCCESS)
    return rc;
  KCHK(adatm_k_rowptr(s, keys, Fd, (uint32_t)nrows, S->rowptr), "rowptr");
  uint32_t const *fids[MAXN] = {0};
  fids[0] = B->root_ids;
  for (uint32_t l = 1; l < d; ++l) fids[l] = g->ids[l];
  for (uint32_t l = 0; l < d; ++l)
    if ((rc = dalloc(e, (void **)&S->rid[l], Fd * 4)) != SPLATT_SUCCESS) return rc;
  KCHK(adatm_k_scatter_rids(s, S->node, Fd, d, (uint32_t const *const *)B->par,
                            fids, (uint32_t *const *)S->rid),
       "scatter rids");
  rc = dwork_build(e, S->rowptr, nrows, e->scat_tile, &S->w, B->stemp, B->tot);
  HIPCHK(hipStreamSynchronize(s), "scatter build sync");
  hipFree(keys);
  hipFree(sortmp);
  return rc;
}

static int engine_add_group_dev(adatm_engine *e, uint32_t gi, dcoo const *c)
{
  uint32_t const N = e->N, R = e->R;
  uint64_t const n = c->nnz;
  hipStream_t const s = e->stream;
  dgroup *g = e->g + gi;
  adatm_group const *pg = e->plan.grp + gi;
  g->memo_mode = pg->memo_mode;
  g->b = pg->begin_imten;
  g->nim = pg->n_imten;
  memcpy(g->order, pg->o_opt, N * sizeof(uint32_t));
  int rc = SPLATT_ERROR_NOMEMORY;
  dbuild B;
  memset(&B, 0, sizeof(B));
  uint32_t *perm = NULL, *keys = NULL, *stemp = NULL;
  uint32_t *sc[MAXN] = {0};
  uint8_t *fd = NULL;
  uint32_t *flags[2] = {0}, *pos[2] = {0}, *start[2] = {0};
  uint64_t const tw = adatm_k_sort_temp_words(n) + adatm_k_scan_temp_words(n);
  if (hipMalloc((void **)&perm, n * 4) || hipMalloc((void **)&keys, n * 4) ||
      hipMalloc((void **)&stemp, tw * 4) || hipMalloc((void **)&fd, n) ||
      hipMalloc((void **)&B.tot, 16))
    goto out;
  B.stemp = stemp;
  for (uint32_t l = 0; l < N; ++l)
    if (hipMalloc((void **)&sc[l], n * 4)) goto out;
  for (int k = 0; k < 2; ++k)
    if (hipMalloc((void **)&flags[k], n * 4) || hipMalloc((void **)&pos[k], n * 4))
      goto out;
  if ((rc = dcoo_lexsort(s, c, g->order, perm, keys, stemp)) != SPLATT_SUCCESS)
    goto out;
  rc = SPLATT_ERROR_DEVICE;
  for (uint32_t l = 0; l < N; ++l)
    if (adatm_k_gather(s, sc[l], c->ind[g->order[l]], perm, n)) goto out;
  if ((rc = dalloc(e, &g->vals, n * e->vsz)) != SPLATT_SUCCESS) goto out;
  rc = SPLATT_ERROR_DEVICE;
  if (adatm_k_gather_vals(s, e->vsz == 8 ? KP_FP64 : KP_FP32, g->vals, c->vals,
                          perm, n))
    goto out;
  if (adatm_k_first_diff(s, (uint32_t const *const *)sc, N, n, fd)) goto out;
  /* levels 0 .. N-2, two at a time for pointers and parents */
  for (uint32_t l = 0; l + 1 < N; ++l) {
    int const cur = l & 1, prv = cur ^ 1;
    uint32_t Fl = 0;
    if (adatm_k_level_flags(s, fd, n, l, flags[cur])) goto out;
    if (dscan(s, flags[cur], pos[cur], n, stemp, B.tot, &Fl) != SPLATT_SUCCESS)
      goto out;
    g->F[l] = Fl;
    uint32_t *fids_l = NULL;
    if (l == 0) {
      if (hipMalloc((void **)&B.root_ids, (size_t)Fl * 4)) goto out;
      fids_l = B.root_ids;
    } else {
      if ((rc = dalloc(e, (void **)&g->ids[l], (size_t)Fl * 4)) != SPLATT_SUCCESS)
        goto out;
      rc = SPLATT_ERROR_DEVICE;
      fids_l = g->ids[l];
    }
    hipFree(start[cur]);
    start[cur] = NULL;
    if (hipMalloc((void **)&start[cur], (size_t)Fl * 4)) goto out;
    if (adatm_k_level_emit(s, flags[cur], pos[cur], sc[l], n, fids_l, start[cur]))
      goto out;
    if (l >= 1) {
      /* fptr of level l-1 (into level l) and parents of level l */
      if ((rc = dalloc(e, (void **)&g->ptr[l - 1], (g->F[l - 1] + 1) * 4)) !=
          SPLATT_SUCCESS)
        goto out;
      rc = SPLATT_ERROR_DEVICE;
      if (adatm_k_fptr(s, start[prv], g->F[l - 1], pos[cur], Fl, g->ptr[l - 1]))
        goto out;
      if (hipMalloc((void **)&B.par[l], (size_t)Fl * 4)) goto out;
      if (adatm_k_parent(s, start[cur], Fl, flags[prv], pos[prv], B.par[l]))
        goto out;
    }
  }
  {
    /* leaf level: nodes are the sorted positions */
    uint32_t const lp = N - 2;
    int const cur = lp & 1;
    g->F[N - 1] = n;
    g->ids[N - 1] = sc[N - 1];
    sc[N - 1] = NULL;
    e->dev_bytes += n * 4;
    if ((rc = dalloc(e, (void **)&g->ptr[lp], (g->F[lp] + 1) * 4)) != SPLATT_SUCCESS)
      goto out;
    rc = SPLATT_ERROR_DEVICE;
    if (adatm_k_fptr(s, start[cur], g->F[lp], NULL, (uint32_t)n, g->ptr[lp])) goto out;
    if (hipMalloc((void **)&B.par[N - 1], n * 4)) goto out;
    if (adatm_k_parent(s, NULL, n, flags[cur], pos[cur], B.par[N - 1])) goto out;
  }
  /* root ids: identity exactly when every root index is occupied */
  g->zero_root = (g->F[0] < e->dims[g->order[0]]);
  if (g->zero_root) {
    g->orow = B.root_ids;
    B.root_ids = NULL;
    e->dev_bytes += g->F[0] * 4;
  }
  /* free the sort's big temporaries before the passes */
  hipFree(keys); keys = NULL;
  hipFree(perm); perm = NULL;
  hipFree(fd); fd = NULL;
  for (uint32_t l = 0; l < N; ++l) {
    hipFree(sc[l]);
    sc[l] = NULL;
  }
  if (!B.root_ids) B.root_ids = g->orow;
  /* memo levels */
  for (uint32_t l = 1; l + 1 < N; ++l)
    if (is_memo_level(g, l) &&
        (rc = dalloc(e, &g->memo[l], g->F[l] * R * e->vsz)) != SPLATT_SUCCESS)
      goto out;
  /* root chain */
  for (uint32_t lb = N - 1; lb > 0;) {
    uint32_t const lt = pass_top(e, g, lb, 0);
    if ((rc = build_pass_dev(e, g, &B, lb, lt, 1)) != SPLATT_SUCCESS) goto out;
    if (lt > 0 && !is_memo_level(g, lt)) {
      size_t const tb = (size_t)g->F[lt] * R * e->vsz;
      if (tb > e->tmp_bytes) e->tmp_bytes = tb;
    }
    lb = lt;
  }
  /* non-root modes */
  for (uint32_t m = 0; m < N; ++m) {
    if (e->use_grp[m] != gi || e->use_tag[m] == 0) continue;
    uint32_t const d = e->level_of[gi][m];
    uint32_t lb = (e->use_tag[m] == 1) ? reuse_source(g, d) : N - 1;
    while (lb > d) {
      uint32_t const lt = lb > d + ADATM_KMAXFUSE ? lb - ADATM_KMAXFUSE : d;
      if ((rc = build_pass_dev(e, g, &B, lb, lt, 0)) != SPLATT_SUCCESS) goto out;
      size_t const tb = (size_t)g->F[lt] * R * e->vsz;
      if (tb > e->tmp_bytes) e->tmp_bytes = tb;
      lb = lt;
    }
    if ((rc = build_scatter_dev(e, g, &B, d)) != SPLATT_SUCCESS) goto out;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    rc = hip_fail(hipGetLastError(), "group build");
    goto out;
  }
  rc = SPLATT_SUCCESS;
out:
  hipStreamSynchronize(s);
  hipFree(perm); hipFree(keys); hipFree(stemp); hipFree(fd); hipFree(B.tot);
  for (uint32_t l = 0; l < N; ++l) hipFree(sc[l]);
  for (int k = 0; k < 2; ++k) {
    hipFree(flags[k]);
    hipFree(pos[k]);
    hipFree(start[k]);
  }
  for (uint32_t l = 0; l < MAXN; ++l) hipFree(B.par[l]);
  if (B.root_ids != g->orow) hipFree(B.root_ids);
  if (rc != SPLATT_SUCCESS && !adatm_last_error()[0])
    adatm_set_error("device group build failed (%d)", rc);
  return rc;
}

int adatm_engine_create(adatm_coo const *t, uint64_t rank,
                        adatm_plan const *plan, int precision, int device,
                        adatm_engine **out)
{
  return adatm_engine_create_tiled(t, rank, plan, precision, device, 0, 0, out);
}

static int engine_create_dev(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                             int precision, int device, uint32_t red_tile,
                             uint32_t scat_tile, uint32_t flags, uint32_t const *use_grp,
                             uint32_t const *use_tag, adatm_engine **out);

int adatm_engine_create_tiled(adatm_coo const *t, uint64_t rank,
                              adatm_plan const *plan, int precision, int device,
                              uint32_t red_tile, uint32_t scat_tile,
                              adatm_engine **out)
{
  return engine_create_dev(t, rank, plan, precision, device, red_tile, scat_tile, 0, NULL,
                           NULL, out);
}

int adatm_engine_create_ex(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                           int precision, int device, uint32_t red_tile,
                           uint32_t scat_tile, uint32_t flags, adatm_engine **out)
{
  uint32_t const known = ADATM_ENGINE_STAGED | ADATM_ENGINE_GLOBAL_NODE_ROWS |
                         (ADATM_ENGINE_STAGED_GROUP(ADATM_MAX_NMODES) -
                          ADATM_ENGINE_STAGED_GROUP(0));
  if (flags & ~known) {
    adatm_set_error("engine: unknown creation flags 0x%x", flags);
    return SPLATT_ERROR_BADINPUT;
  }
  return engine_create_dev(t, rank, plan, precision, device, red_tile, scat_tile, flags,
                           NULL, NULL, out);
}

/* built on the device with the caller's mode -> (group, tag) map (the SPLATT
 * layer's plain and space-efficient CSF sets) */
int adatm_engine_create_use(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                            int precision, int device, uint32_t const *use_grp,
                            uint32_t const *use_tag, adatm_engine **out)
{
  return engine_create_dev(t, rank, plan, precision, device, 0, 0, 0, use_grp, use_tag,
                           out);
}

static int engine_create_dev(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                             int precision, int device, uint32_t red_tile,
                             uint32_t scat_tile, uint32_t flags, uint32_t const *use_grp,
                             uint32_t const *use_tag, adatm_engine **out)
{
  *out = NULL;
  uint32_t const N = t->nmodes;
  if (N < 3 || N > MAXN || t->nnz == 0 || t->nnz >= (1ull << 32))
    return SPLATT_ERROR_BADINPUT;
  if ((red_tile && (red_tile < 8 || red_tile > 4096)) ||
      (scat_tile && (scat_tile < 8 || scat_tile > 4096))) {
    adatm_set_error("engine: work-item tiles must be in [8, 4096]");
    return SPLATT_ERROR_BADINPUT;
  }
  adatm_set_error("%s", "");
  adatm_engine *e = (adatm_engine *)calloc(1, sizeof(adatm_engine));
  if (!e) return SPLATT_ERROR_NOMEMORY;
  e->red_tile = red_tile;
  e->scat_tile = scat_tile;
  e->flags = flags;
  int rc = engine_begin(e, N, t->dims, t->nnz, rank, plan, precision, device,
                        use_grp, use_tag);
  dcoo c;
  memset(&c, 0, sizeof(c));
  if (rc == SPLATT_SUCCESS) rc = dcoo_upload(&c, t, 1);
  for (uint32_t gi = 0; rc == SPLATT_SUCCESS && gi < plan->n_grp; ++gi)
    rc = engine_add_group_dev(e, gi, &c);
  dcoo_free(&c);
  if (rc == SPLATT_SUCCESS) rc = engine_finish(e);
  e->ttnormsq = adatm_sumsq(t->vals, t->nnz);
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
      hipFree(g->cfd[l]);
      hipFree(g->memo[l]);
      dscat *S = g->scat + l;
      hipFree(S->node);
      hipFree(S->rowptr);
      for (uint32_t k = 0; k < ADATM_KMAXSCAT; ++k) hipFree(S->rid[k]);
      free_work(&S->w);
    }
    for (uint32_t i = 0; i < g->npass; ++i) {
      dpass *P = g->pass + i;
      free_work(&P->w);
      for (uint32_t j = 0; j <= ADATM_KMAXFUSE; ++j) free_work(&P->sw[j]);
      hipFree(P->shead);
      hipFree(P->stail);
    }
    hipFree(g->orow);
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
  hipFree(e->carry2);
  hipFree(e->spart);
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
    ev_begin(e, KC_DENSE, 0, 0.0);
    KCHK(adatm_k_solve_inv(e->stream, e->G, N, m, R, e->Vinv, e->status),
         "solve_inv");
    KCHK(adatm_k_update(e->stream, e->prec, e->M, e->Vinv, e->A[m], e->Abf[m],
                        e->dims[m], R, e->iter, e->lambda, e->part,
                        e->G + (size_t)m * R * R),
         "update");
    ev_end(e);
  }
  ev_begin(e, KC_DENSE, 0, 0.0);
  KCHK(adatm_k_fit(e->stream, e->prec, e->A[N - 1], e->M, e->dims[N - 1],
                   e->lambda, e->G, N, R, e->ttnormsq, e->part, e->fits,
                   e->fits_cap, e->iter),
       "fit");
  ev_end(e);
  return SPLATT_SUCCESS;
}

/* Grams of the current factors, the sweep counter and the solve status
 * reset */
static int cpd_prepare(adatm_engine *e)
{
  uint32_t const R = e->R;
  for (uint32_t m = 0; m < e->N; ++m)
    KCHK(adatm_k_gram(e->stream, e->prec, e->A[m], e->dims[m], R, e->part,
                      e->G + (size_t)m * R * R),
         "gram");
  KCHK(adatm_k_zero(e->stream, e->iter, 4 * sizeof(uint32_t)), "zero iter");
  KCHK(adatm_k_zero(e->stream, e->status, 4 * sizeof(int)), "zero status");
  e->nsweep = 0;
  return SPLATT_SUCCESS;
}

/* worst solve status since cpd_prepare (the stream is synchronised here) */
static int read_status(adatm_engine *e, int *st)
{
  HIPCHK(hipMemcpyAsync(st, e->status, sizeof(int), hipMemcpyDeviceToHost, e->stream),
         "status download");
  HIPCHK(hipStreamSynchronize(e->stream), "status sync");
  return SPLATT_SUCCESS;
}

/* SPLATT_ERROR_NUMERIC when a solve found no positive pivot even with the
 * largest ridge (the factors and fits are then not finite) */
static int status_rc(int st)
{
  if (st >= 2) {
    adatm_set_error("CP-ALS: the normal equations (hadamard of Grams) had no positive "
                    "pivot even with a ridge of 1e-6 of their largest diagonal");
    return SPLATT_ERROR_NUMERIC;
  }
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

/* one ALS sweep; its fit goes to fits[nsweep], so no sweep is launched once
 * the buffer is full (adatm_engine_cpd_begin starts over) */
static int sweep_launch(adatm_engine *e, int use_graph)
{
  if (e->nsweep >= e->fits_cap) {
    adatm_set_error("engine: %u sweeps since the last CP-ALS start fill the fits "
                    "buffer; start again (adatm_engine_cpd_begin)", e->nsweep);
    return SPLATT_ERROR_BADINPUT;
  }
  if (use_graph) {
    int const rc = graph_build(e);
    if (rc != SPLATT_SUCCESS) return rc;
    HIPCHK(hipGraphLaunch(e->gexec, e->stream), "graph launch");
  } else {
    int const rc = sweep_enqueue(e);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  e->nsweep++;
  return SPLATT_SUCCESS;
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
  int st = 0;
  rc = read_status(e, &st);
  return rc != SPLATT_SUCCESS ? rc : status_rc(st);
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
  double t[1000]; /* nrep <= 1000, checked above */
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a), "event");
  hipError_t const eb = hipEventCreate(&b);
  if (eb != hipSuccess) {
    hipEventDestroy(a);
    return hip_fail(eb, "event");
  }
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
  if (rc != SPLATT_SUCCESS) return rc;
  qsort(t, (size_t)nrep, sizeof(double), cmp_double);
  *ms_median = (nrep % 2) ? t[nrep / 2] : 0.5 * (t[nrep / 2 - 1] + t[nrep / 2]);

  if (per_kernel_ms) {
    /* one instrumented, non-graph pass: per-class device time and bytes */
    double tsum[KC_NCLASS] = {0}, bsum[KC_NCLASS] = {0};
    int cnt[KC_NCLASS] = {0};
    if (what == 1 && e->nsweep >= e->fits_cap) {
      rc = cpd_prepare(e); /* the pass writes fits[nsweep] */
      if (rc != SPLATT_SUCCESS) return rc;
    }
    e->instrument = 1;
    e->nev = 0;
    rc = (what == 0) ? seq_enqueue(e) : sweep_enqueue(e);
    e->instrument = 0;
    if (rc != SPLATT_SUCCESS) return rc;
    if (what == 1) e->nsweep++;
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

/* per-launch records of the last instrumented pass of adatm_engine_time:
 * class, levels fused / scatter level, device ms and algorithmic bytes of
 * every launch, in order */
int adatm_engine_launches(adatm_engine *e, int max, int *cls, int *info,
                          double *ms, double *bytes)
{
  if (!e || max < 0) return -1;
  int const n = e->nev < max ? e->nev : max;
  for (int i = 0; i < n; ++i) {
    float t = 0.f;
    hipEventElapsedTime(&t, e->ev[2 * i], e->ev[2 * i + 1]);
    if (cls) cls[i] = e->ev_class[i];
    if (info) info[i] = e->ev_info[i];
    if (ms) ms[i] = t;
    if (bytes) bytes[i] = e->ev_bytes[i];
  }
  return n;
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
    dgroup *g = e->g + e->use_grp[m];
    uint32_t lb, stop;
    if (e->use_tag[m] == 0) {
      lb = N - 1;
      stop = 0;
    } else {
      stop = e->level_of[e->use_grp[m]][m];
      lb = (e->use_tag[m] == 1) ? reuse_source(g, stop) : N - 1;
    }
    while (lb > stop) {
      uint32_t const lt = (e->use_tag[m] == 0)
                              ? pass_top(e, g, lb, 0)
                              : (lb > stop + ADATM_KMAXFUSE
                                     ? lb - ADATM_KMAXFUSE : stop);
      B += bytes_pass(e, g, lb, lt);
      for (uint32_t l = lt + 1; l <= lb; ++l) Fl += 2.0 * R * (double)g->F[l];
      lb = lt;
    }
    if (e->use_tag[m] != 0) {
      B += bytes_scatter(e, g, stop);
      Fl += R * (double)(stop + 1) * (double)g->F[stop];
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

/* Time nsteps back-to-back repetitions (what = 0 MTTKRP sequence, 1 ALS
 * sweep) between two events on the engine's stream, synchronised before and
 * after; *ms_total is the elapsed device time. */
int adatm_engine_run_timed(adatm_engine *e, int what, uint32_t nsteps,
                           int use_graph, double *ms_total)
{
  if (!e || nsteps == 0) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int rc;
  if (what == 1) {
    if (nsteps >= e->fits_cap) {
      adatm_set_error("engine: at most %u timed sweeps per call", e->fits_cap - 1);
      return SPLATT_ERROR_BADINPUT;
    }
    rc = cpd_prepare(e);
    if (rc != SPLATT_SUCCESS) return rc;
    if (use_graph) {
      rc = graph_build(e);
      if (rc != SPLATT_SUCCESS) return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(e->stream), "pre sync");
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a), "event");
  hipError_t const eb = hipEventCreate(&b);
  if (eb != hipSuccess) {
    hipEventDestroy(a);
    return hip_fail(eb, "event");
  }
  hipEventRecord(a, e->stream);
  rc = SPLATT_SUCCESS;
  for (uint32_t i = 0; i < nsteps && rc == SPLATT_SUCCESS; ++i)
    rc = (what == 0) ? seq_enqueue(e) : sweep_launch(e, use_graph);
  hipEventRecord(b, e->stream);
  hipError_t const s = hipEventSynchronize(b);
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != SPLATT_SUCCESS) return rc;
  HIPCHK(s, "timed sync");
  *ms_total = ms;
  return SPLATT_SUCCESS;
}

/* fits of the sweeps run since the last CP-ALS preparation */
int adatm_engine_get_fits(adatm_engine *e, uint32_t n, double *fits)
{
  if (!e || n > e->fits_cap) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(e->stream), "sync");
  HIPCHK(hipMemcpy(fits, e->fits, n * sizeof(double), hipMemcpyDeviceToHost),
         "fits download");
  return SPLATT_SUCCESS;
}

/* Stepwise CP-ALS for callers that report every sweep: begin computes the
 * Grams of the current factors and resets the sweep counter; each sweep runs
 * synchronously and returns its fit. */
int adatm_engine_cpd_begin(adatm_engine *e)
{
  if (!e) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int const rc = cpd_prepare(e);
  if (rc != SPLATT_SUCCESS) return rc;
  HIPCHK(hipStreamSynchronize(e->stream), "cpd begin sync");
  return SPLATT_SUCCESS;
}

/* `it` must be the number of sweeps since adatm_engine_cpd_begin: the fit
 * returned is the one this sweep writes (fits[it]) */
int adatm_engine_cpd_sweep(adatm_engine *e, uint32_t it, int use_graph,
                           double *fit)
{
  if (!e || !fit) return SPLATT_ERROR_BADINPUT;
  if (it != e->nsweep) {
    adatm_set_error("engine: sweep %u requested, %u sweeps run since the last "
                    "CP-ALS start", it, e->nsweep);
    return SPLATT_ERROR_BADINPUT;
  }
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  int rc = sweep_launch(e, use_graph);
  if (rc != SPLATT_SUCCESS) return rc;
  HIPCHK(hipMemcpyAsync(fit, e->fits + (e->nsweep - 1), sizeof(double),
                        hipMemcpyDeviceToHost, e->stream),
         "fit download");
  int st = 0;
  rc = read_status(e, &st); /* synchronises the stream */
  return rc != SPLATT_SUCCESS ? rc : status_rc(st);
}

int adatm_engine_solve_status(adatm_engine *e, int *status)
{
  if (!e || !status) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(e->device), "hipSetDevice");
  return read_status(e, status);
}

uint64_t adatm_engine_memo_bytes(adatm_engine const *e, uint32_t grp,
                                 uint32_t level)
{
  if (!e || grp >= e->plan.n_grp || level >= e->N) return 0;
  return e->g[grp].memo[level] ? e->g[grp].F[level] * e->R * e->vsz : 0;
}

int adatm_engine_set_batch(adatm_engine *e, uint32_t batch)
{
  uint32_t const rb = batch & ~(ADATM_BATCH_PIPELINED | ADATM_BATCH_WIDE);
  if (!e || (rb != 0 && rb != 4 && rb != 8 && rb != 16))
    return SPLATT_ERROR_BADINPUT;
  if (batch != e->batch && e->gvalid) { /* the captured sweep holds the old one */
    hipGraphExecDestroy(e->gexec);
    hipGraphDestroy(e->graph);
    e->gvalid = 0;
  }
  e->batch = batch;
  return SPLATT_SUCCESS;
}

/* Work-item tiles, root-pass form and reduction batch of one plan chosen by
 * timing: the reduction tile over {32, 64, 128, 256, 512} with the scatter
 * tile at its default, then the scatter tile over {16, 32, 64, 128, 256}
 * with the best reduction tile, then (when the scatter tile moved) the
 * reduction tile again with the new scatter tile; then, with the best
 * tiles, every memo group's root pass staged (against the default, one
 * pass to the root), then each group's form flipped in turn, keeping what
 * is faster; then the
 * default batch against the children per batch over {4, 8, 16}, the
 * pipelined 4 and 8 and the wide 4 and 8.  The defaults are timed first and
 * a candidate displaces the best so far only when it is more than
 * TUNE_MARGIN faster: near-equal choices differ by run-to-run noise, and
 * without the margin the choice (and with it the kernel that dominates the
 * sweep) flips between runs.  The order of the calls to `timer` is part of
 * the contract (the unit tests drive it with a table of timings). */
#define TUNE_MARGIN 0.99
int adatm_autotune_search(adatm_tile_timer timer, void *ctx, uint32_t groups,
                          uint32_t *red_tile, uint32_t *scat_tile, uint32_t *batch,
                          uint32_t *flags, double *best_ms)
{
  if (!timer || !red_tile || !scat_tile) return SPLATT_ERROR_BADINPUT;
  /* defaults first */
  static uint32_t const RT[5] = {RED_TILE, 32, 64, 256, 512},
                        ST[5] = {SCAT_TILE, 16, 32, 128, 256};
  uint32_t br = RED_TILE, bs = SCAT_TILE;
  double bt = INFINITY;
  uint32_t r1 = 0; /* the reduction tile after the second pass */
  for (int pass = 0; pass < 3; ++pass) {
    if (pass == 2) {
      if (bs == SCAT_TILE) break; /* every reduction tile was timed with it */
      r1 = br;
    }
    for (int i = 0; i < 5; ++i) {
      uint32_t const r = pass == 1 ? br : RT[i];
      uint32_t const sc = pass == 0 ? SCAT_TILE : (pass == 1 ? ST[i] : bs);
      if (pass == 1 && sc == SCAT_TILE) continue; /* timed in the first pass */
      if (pass == 2 && r == r1) continue;         /* timed in the second pass */
      double ms = INFINITY;
      int const rc = timer(ctx, r, sc, 0, 0, &ms);
      if (rc != SPLATT_SUCCESS) return rc;
      if (ms < bt * TUNE_MARGIN) {
        bt = ms;
        br = r;
        bs = sc;
      }
    }
  }
  /* root passes: every memo group staged, then each group's form flipped in
     turn from the better of the two (keeping what helped) */
  uint32_t bf = 0, all = 0;
  for (uint32_t g = 0; g < ADATM_MAX_NMODES; ++g)
    if ((groups >> g) & 1u) all |= ADATM_ENGINE_STAGED_GROUP(g);
  for (int k = -1; all && k < ADATM_MAX_NMODES; ++k) {
    if (k >= 0 && !((groups >> k) & 1u)) continue;
    uint32_t const f = (k < 0) ? all : bf ^ ADATM_ENGINE_STAGED_GROUP(k);
    double ms = INFINITY;
    int const rc = timer(ctx, br, bs, 0, f, &ms);
    if (rc != SPLATT_SUCCESS) return rc;
    if (ms < bt * TUNE_MARGIN) {
      bt = ms;
      bf = f;
    }
  }
  static uint32_t const BT[7] = {4, 8, 16, 4 | ADATM_BATCH_PIPELINED,
                                 8 | ADATM_BATCH_PIPELINED, 4 | ADATM_BATCH_WIDE,
                                 8 | ADATM_BATCH_WIDE};
  uint32_t bb = 0;
  for (int i = 0; i < 7; ++i) {
    double ms = INFINITY;
    int const rc = timer(ctx, br, bs, BT[i], bf, &ms);
    if (rc != SPLATT_SUCCESS) return rc;
    if (ms < bt * TUNE_MARGIN) {
      bt = ms;
      bb = BT[i];
    }
  }
  *red_tile = br;
  *scat_tile = bs;
  if (batch) *batch = bb;
  if (flags) *flags = bf;
  if (best_ms) *best_ms = bt;
  return SPLATT_SUCCESS;
}

/* the device timer of adatm_autotune_tiles: CP-ALS sweeps of an engine
 * built with the tiles and flags (kept while only the batch changes) */
typedef struct {
  adatm_coo const *t;
  uint64_t rank;
  adatm_plan const *plan;
  int precision, device, nrep;
  adatm_engine *e;
  uint32_t red, scat, flags;
} tune_ctx;

static int tune_time(void *vctx, uint32_t red, uint32_t scat, uint32_t batch,
                     uint32_t flags, double *ms)
{
  tune_ctx *c = (tune_ctx *)vctx;
  if (!c->e || c->red != red || c->scat != scat || c->flags != flags) {
    adatm_engine_destroy(c->e);
    c->e = NULL;
    int rc = adatm_engine_create_ex(c->t, c->rank, c->plan, c->precision, c->device, red,
                                    scat, flags, &c->e);
    if (rc == SPLATT_SUCCESS) rc = adatm_engine_init_factors(c->e, 1);
    if (rc != SPLATT_SUCCESS) return rc;
    c->red = red;
    c->scat = scat;
    c->flags = flags;
  }
  int const rc = adatm_engine_set_batch(c->e, batch);
  if (rc != SPLATT_SUCCESS) return rc;
  return adatm_engine_time(c->e, 1, 1, c->nrep, 1, ms, NULL);
}

int adatm_autotune_tiles(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                         int precision, int device, int nrep, uint32_t *red_tile,
                         uint32_t *scat_tile, uint32_t *batch, uint32_t *flags,
                         double *best_ms)
{
  tune_ctx c;
  memset(&c, 0, sizeof(c));
  c.t = t;
  c.rank = rank;
  c.plan = plan;
  c.precision = precision;
  c.device = device;
  c.nrep = nrep;
  /* the memo groups' root passes may be fused or staged; others have one form */
  uint32_t groups = 0;
  for (uint32_t g = 0; g < plan->n_grp; ++g)
    if (plan->grp[g].n_imten > 0) groups |= 1u << g;
  int const rc = adatm_autotune_search(tune_time, &c, groups, red_tile, scat_tile, batch,
                                       flags, best_ms);
  adatm_engine_destroy(c.e);
  return rc;
}

/* ======================================================================== *
 * Batched CP-ALS: many independent engines on one device, every instance's
 * ALS sweep captured in one hipGraph.  Each instance runs on a branch stream
 * of the batch (not the engine's own stream, which never joins the batch's
 * capture), forked from and joined to the batch stream inside the capture,
 * so the graph's branches are independent and small instances run side by
 * side.
 * ======================================================================== */
/* one round of the grouped sweep: every instance with more than r modes runs
 * its mode-r MTTKRP on its own stream, then one grouped dense update */
typedef struct {
  uint32_t n, desc0, Rmax;            /* instances, first descriptor, largest R */
  uint32_t nb_mm, nb_gr, nb_sp;       /* grouped grids */
  uint32_t mm_off, gr_off, sp_off;    /* their block maps in dmap */
} batch_round;

typedef struct {
  adatm_kstage kind; /* type, leaf, K, side, d, zs of the launch's stages */
  uint64_t bmap0;    /* its first block-map entry */
  uint32_t nb;       /* its blocks */
} glaunch;

struct adatm_batch {
  int n, device;
  adatm_engine **e;
  hipStream_t s;
  hipStream_t *bs; /* branch stream of each instance */
  hipEvent_t fork, *join;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int gvalid;
  /* grouped dense updates (all engines of one precision) */
  int grouped;
  uint32_t nround;
  batch_round *round;
  uint32_t fit0, nb_in, in_off; /* the fits: descriptors, grid, block map */
  adatm_kdense *ddesc; /* device descriptors of every round, then the fits */
  uint32_t *dmap;      /* device block maps of every round, then the fits' */
  /* the form of the MTTKRPs of a round (ADATM_FORM_*): wanted, last built */
  int want_form, form;
  /* grouped form: per round, one launch per depth and kind of stage */
  adatm_kstage *gst;   /* device stages of every round */
  uint64_t *gbmap;     /* device block maps of every launch */
  glaunch *gl;         /* the launches of every round, round by round */
  uint32_t *gl0;       /* first launch of each round (nround + 1) */
  /* persistent form: each round's MTTKRPs in one adatm_k_persist launch */
  uint32_t nq, pgrid, spin;
  uint32_t xmap[16];
  uint32_t *p_stage0, *p_nstage, *p_task0; /* per round, into the arrays below */
  adatm_kstage *dstages;
  uint64_t *dtasks;
  uint32_t *dqoff; /* (nq + 1) per round */
  uint32_t *dxmap, *dctr, *dstat;
  uint32_t ctr_words;
  uint64_t *p_served; /* per queue: tasks of one sweep */
};

void adatm_batch_destroy(adatm_batch *b)
{
  if (!b) return;
  if (b->gvalid) {
    hipGraphExecDestroy(b->gexec);
    hipGraphDestroy(b->graph);
  }
  if (b->join)
    for (int i = 0; i < b->n; ++i)
      if (b->join[i]) hipEventDestroy(b->join[i]);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->bs)
    for (int i = 0; i < b->n; ++i)
      if (b->bs[i]) hipStreamDestroy(b->bs[i]);
  free(b->bs);
  if (b->s) hipStreamDestroy(b->s);
  if (b->ddesc) hipFree(b->ddesc);
  if (b->dmap) hipFree(b->dmap);
  hipFree(b->dstages);
  hipFree(b->dtasks);
  hipFree(b->dqoff);
  hipFree(b->dxmap);
  hipFree(b->dctr);
  hipFree(b->dstat);
  free(b->p_stage0);
  free(b->p_nstage);
  free(b->p_task0);
  free(b->p_served);
  hipFree(b->gst);
  hipFree(b->gbmap);
  free(b->gl);
  free(b->gl0);
  free(b->round);
  free(b->join);
  free(b->e);
  free(b);
}

/* Descriptors and block maps of the grouped sweep, uploaded once: round r
 * holds every instance with more than r modes at mode r, with the block
 * counts of its own single-instance launches. */
static int batch_rounds(adatm_batch *b)
{
  uint32_t nround = 0, ndesc = 0;
  for (int i = 0; i < b->n; ++i)
    if (b->e[i]->N > nround) nround = b->e[i]->N;
  for (int i = 0; i < b->n; ++i) ndesc += b->e[i]->N + 1; /* + its fit */
  free(b->round); /* a failed earlier attempt */
  b->round = (batch_round *)calloc(nround, sizeof(batch_round));
  adatm_kdense *hd = (adatm_kdense *)calloc(ndesc, sizeof(adatm_kdense));
  uint64_t nmap = 0;
  for (int i = 0; i < b->n; ++i)
    for (uint32_t m = 0; m < b->e[i]->N; ++m) {
      uint32_t a, g, p;
      adatm_k_dense_blocks(b->e[i]->dims[m], b->e[i]->R, &a, &g, &p);
      nmap += (uint64_t)a + g + p;
    }
  for (int i = 0; i < b->n; ++i) nmap += adatm_k_fit_blocks(b->e[i]->dims[b->e[i]->N - 1]);
  uint32_t *hm = (uint32_t *)malloc((size_t)nmap * sizeof(uint32_t));
  if (!b->round || !hd || !hm) {
    free(hd);
    free(hm);
    return SPLATT_ERROR_NOMEMORY;
  }
  uint32_t nd = 0, nm = 0;
  for (uint32_t r = 0; r < nround; ++r) {
    batch_round *B = b->round + r;
    B->desc0 = nd;
    for (int i = 0; i < b->n; ++i) {
      adatm_engine *e = b->e[i];
      if (e->N <= r) continue;
      adatm_kdense *x = hd + nd;
      x->N = e->N;
      x->m = r;
      x->R = e->R;
      x->I = e->dims[r];
      x->G = e->G;
      x->Vinv = e->Vinv;
      x->status = e->status;
      x->M = e->M;
      x->A = e->A[r];
      x->Abf = e->Abf[r];
      x->iter = e->iter;
      x->lambda = e->lambda;
      x->part = e->part;
      x->G_out = e->G + (size_t)r * e->R * e->R;
      adatm_k_dense_blocks(x->I, x->R, &x->nmm, &x->ngr, &x->nsp);
      x->mm0 = B->nb_mm;
      x->gr0 = B->nb_gr;
      x->sp0 = B->nb_sp;
      B->nb_mm += x->nmm;
      B->nb_gr += x->ngr;
      B->nb_sp += x->nsp;
      if (e->R > B->Rmax) B->Rmax = e->R;
      ++B->n;
      ++nd;
    }
    uint32_t *off[3] = {&B->mm_off, &B->gr_off, &B->sp_off};
    for (int k = 0; k < 3; ++k) {
      *off[k] = nm;
      for (uint32_t q = 0; q < B->n; ++q) {
        adatm_kdense const *x = hd + B->desc0 + q;
        uint32_t const c = k == 0 ? x->nmm : (k == 1 ? x->ngr : x->nsp);
        for (uint32_t t = 0; t < c; ++t) hm[nm++] = q;
      }
    }
  }
  /* the fits: every instance at its last mode */
  b->fit0 = nd;
  b->nb_in = 0;
  for (int i = 0; i < b->n; ++i) {
    adatm_engine *e = b->e[i];
    adatm_kdense *x = hd + nd++;
    x->N = e->N;
    x->m = e->N - 1;
    x->R = e->R;
    x->I = e->dims[e->N - 1];
    x->G = e->G;
    x->A = e->A[e->N - 1];
    x->M = e->M;
    x->lambda = e->lambda;
    x->part = e->part;
    x->iter = e->iter;
    x->ttnormsq = e->ttnormsq;
    x->fits = e->fits;
    x->fcap = e->fits_cap;
    x->nin = adatm_k_fit_blocks(x->I);
    x->in0 = b->nb_in;
    b->nb_in += x->nin;
  }
  b->in_off = nm;
  for (int i = 0; i < b->n; ++i)
    for (uint32_t t = 0; t < hd[b->fit0 + i].nin; ++t) hm[nm++] = (uint32_t)i;
  b->nround = nround;
  int rc = SPLATT_SUCCESS;
  if (hipMalloc((void **)&b->ddesc, (size_t)nd * sizeof(adatm_kdense)) != hipSuccess ||
      hipMalloc((void **)&b->dmap, (size_t)(nm ? nm : 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMemcpy(b->ddesc, hd, (size_t)nd * sizeof(adatm_kdense), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->dmap, hm, (size_t)nm * sizeof(uint32_t), hipMemcpyHostToDevice) !=
          hipSuccess)
    rc = hip_fail(hipGetLastError(), "batch descriptors");
  if (rc != SPLATT_SUCCESS) {
    hipFree(b->ddesc);
    hipFree(b->dmap);
    b->ddesc = NULL;
    b->dmap = NULL;
  }
  free(hd);
  free(hm);
  return rc;
}

/* instance i's work (its MTTKRP of mode `mode`, or with mode < 0 its whole
 * sweep) enqueued on its branch stream, forked from and joined to the batch
 * stream; the engine's launches go to the branch stream while enqueued */
static int batch_branch(adatm_batch *b, int i, int mode)
{
  adatm_engine *e = b->e[i];
  if (hipStreamWaitEvent(b->bs[i], b->fork, 0) != hipSuccess) return SPLATT_ERROR_DEVICE;
  hipStream_t const own = e->stream;
  e->stream = b->bs[i];
  int const rc = (mode < 0) ? sweep_enqueue(e) : mttkrp_dev(e, (uint32_t)mode);
  e->stream = own;
  if (rc != SPLATT_SUCCESS) return rc;
  if (hipEventRecord(b->join[i], b->bs[i]) != hipSuccess ||
      hipStreamWaitEvent(b->s, b->join[i], 0) != hipSuccess)
    return SPLATT_ERROR_DEVICE;
  return SPLATT_SUCCESS;
}

/* ---- persistent mode -------------------------------------------------------
 * Each round's MTTKRPs run in one adatm_k_persist launch.  Every instance's
 * launch sequence for its mode-r MTTKRP is recorded as a chain of stages
 * (each waiting on the one before), the chains are spread over one queue
 * per XCD by their block counts (largest first onto the least-loaded
 * queue), and a queue lists its tasks depth by depth: every instance's
 * first stage, then every second stage, ..., so a stage's dependency always
 * comes earlier in its queue. */

#define PERSIST_SPIN (1u << 22) /* polls of one dependency wait (s_sleep 2 each) */

int adatm_persist_plan(uint32_t n, uint32_t const *len, double const *cost, uint32_t nq,
                       uint32_t *queue, uint32_t *order, uint32_t *qstart)
{
  if ((n && (!len || !cost || !queue)) || nq == 0 || !order || !qstart)
    return SPLATT_ERROR_BADINPUT;
  uint32_t *first = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *ord = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  double *work = (double *)malloc((n ? n : 1) * sizeof(double));
  double *load = (double *)malloc(nq * sizeof(double));
  if (!first || !ord || !work || !load) {
    free(first);
    free(ord);
    free(work);
    free(load);
    return SPLATT_ERROR_NOMEMORY;
  }
  uint32_t tot = 0, maxlen = 0;
  for (uint32_t i = 0; i < n; ++i) {
    first[i] = tot;
    work[i] = 0.0;
    for (uint32_t d = 0; d < len[i]; ++d) work[i] += cost[tot + d];
    tot += len[i];
    if (len[i] > maxlen) maxlen = len[i];
    ord[i] = i;
  }
  /* largest chains first (ties: lower index), each onto the least-loaded
     queue (ties: lower queue) */
  for (uint32_t i = 1; i < n; ++i) {
    uint32_t const v = ord[i];
    uint32_t j = i;
    while (j > 0 && work[ord[j - 1]] < work[v]) {
      ord[j] = ord[j - 1];
      --j;
    }
    ord[j] = v;
  }
  for (uint32_t q = 0; q < nq; ++q) load[q] = 0.0;
  for (uint32_t t = 0; t < n; ++t) {
    uint32_t const i = ord[t];
    uint32_t best = 0;
    for (uint32_t q = 1; q < nq; ++q)
      if (load[q] < load[best]) best = q;
    queue[i] = best;
    load[best] += work[i];
  }
  /* queue by queue, depth by depth, the chains of a depth in that order */
  uint32_t m = 0;
  for (uint32_t q = 0; q < nq; ++q) {
    qstart[q] = m;
    for (uint32_t d = 0; d < maxlen; ++d)
      for (uint32_t t = 0; t < n; ++t) {
        uint32_t const i = ord[t];
        if (queue[i] == q && d < len[i]) order[m++] = first[i] + d;
      }
  }
  qstart[nq] = m;
  free(first);
  free(ord);
  free(work);
  free(load);
  return SPLATT_SUCCESS;
}

/* the XCD ids the device's workgroups report: xmap[id] = queue (~0: none) */
static int batch_probe_xcc(adatm_batch *b)
{
  uint32_t const nb = 4096;
  uint32_t *d = NULL, *h = (uint32_t *)malloc(nb * sizeof(uint32_t));
  if (!h) return SPLATT_ERROR_NOMEMORY;
  int rc = SPLATT_SUCCESS;
  if (hipMalloc((void **)&d, nb * sizeof(uint32_t)) != hipSuccess ||
      adatm_k_xcc_probe(b->s, d, nb) != 0 ||
      hipMemcpyAsync(h, d, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, b->s) != hipSuccess ||
      hipStreamSynchronize(b->s) != hipSuccess)
    rc = hip_fail(hipGetLastError(), "XCD probe");
  hipFree(d);
  if (rc == SPLATT_SUCCESS) {
    int seen[16] = {0};
    for (uint32_t i = 0; i < nb; ++i) seen[h[i] & 15u] = 1;
    b->nq = 0;
    for (int x = 0; x < 16; ++x) b->xmap[x] = seen[x] ? b->nq++ : 0xFFFFFFFFu;
  }
  free(h);
  return rc;
}

/* instances eligible for the persistent launch: one precision (grouped),
 * 8 <= R <= 32 and at most 6 modes (stages of <= 5 fused levels) */
static int batch_persist_ok(adatm_batch const *b)
{
  if (!b->grouped) return 0;
  for (int i = 0; i < b->n; ++i)
    if (b->e[i]->R < 8 || b->e[i]->R > 32 || b->e[i]->N > 6) return 0;
  return 1;
}

static int batch_record_round(adatm_batch *b, uint32_t r, krec *all, uint32_t s0,
                              uint32_t *first, uint32_t *len);

static int batch_persist_build(adatm_batch *b)
{
  int rc = batch_probe_xcc(b);
  if (rc != SPLATT_SUCCESS) return rc;
  if (b->nq == 0) return SPLATT_ERROR_DEVICE;
  KCHK(adatm_k_persist_grid(b->e[0]->prec, b->device, &b->pgrid), "persistent grid");
  uint32_t const nr = b->nround, nq = b->nq, n = (uint32_t)b->n;
  free(b->p_stage0); /* a failed earlier attempt */
  free(b->p_nstage);
  free(b->p_task0);
  free(b->p_served);
  krec all = {0};
  uint64_t *tasks = NULL;
  uint64_t ntask = 0, task_cap = 0;
  uint32_t *qoff = (uint32_t *)calloc((size_t)nr * (nq + 1), sizeof(uint32_t));
  uint32_t *first = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *len = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *queue = (uint32_t *)malloc(n * sizeof(uint32_t));
  b->p_stage0 = (uint32_t *)calloc(nr, sizeof(uint32_t));
  b->p_nstage = (uint32_t *)calloc(nr, sizeof(uint32_t));
  b->p_task0 = (uint32_t *)calloc(nr, sizeof(uint32_t));
  b->p_served = (uint64_t *)calloc(nq, sizeof(uint64_t));
  if (!qoff || !first || !len || !queue || !b->p_stage0 ||
      !b->p_nstage || !b->p_task0 || !b->p_served)
    rc = SPLATT_ERROR_NOMEMORY;
  b->ctr_words = 0;
  for (uint32_t r = 0; rc == SPLATT_SUCCESS && r < nr; ++r) {
    uint32_t const s0 = all.n;
    b->p_stage0[r] = s0;
    b->p_task0[r] = (uint32_t)ntask;
    /* each instance's chain of stages, recorded as its own launches */
    rc = batch_record_round(b, r, &all, s0, first, len);
    if (rc != SPLATT_SUCCESS) break;
    b->p_nstage[r] = all.n - s0;
    /* chains onto queues and each queue's stage order (adatm_persist_plan:
       depth by depth, so a stage always follows the one it waits for;
       ordering by the share of its chain's cost before a stage measured
       slower on the 64-instance workload, 5.70 against 5.52 ms per sweep) */
    uint32_t const ns_r = all.n - s0;
    uint32_t *sord = (uint32_t *)malloc((ns_r ? ns_r : 1) * sizeof(uint32_t));
    uint32_t *qst = (uint32_t *)malloc((nq + 1) * sizeof(uint32_t));
    if (!sord || !qst) rc = SPLATT_ERROR_NOMEMORY;
    if (rc == SPLATT_SUCCESS)
      rc = adatm_persist_plan(n, len, all.cost + s0, nq, queue, sord, qst);
    uint32_t *qo = qoff + (size_t)r * (nq + 1);
    uint64_t const t0 = ntask;
    for (uint32_t q = 0; rc == SPLATT_SUCCESS && q < nq; ++q) {
      qo[q] = (uint32_t)(ntask - t0);
      for (uint32_t a = qst[q]; rc == SPLATT_SUCCESS && a < qst[q + 1]; ++a) {
        adatm_kstage const *x = all.st + s0 + sord[a];
        if (ntask + x->nblocks > task_cap) {
          uint64_t cap = task_cap ? 2 * task_cap : 4096;
          while (cap < ntask + x->nblocks) cap *= 2;
          uint64_t *nt = (uint64_t *)realloc(tasks, cap * sizeof(uint64_t));
          if (!nt) {
            rc = SPLATT_ERROR_NOMEMORY;
            continue;
          }
          tasks = nt;
          task_cap = cap;
        }
        uint64_t const sl = (uint64_t)sord[a];
        for (uint32_t v = 0; v < x->nblocks; ++v) tasks[ntask++] = (sl << 32) | v;
      }
      b->p_served[q] += (ntask - t0) - qo[q];
    }
    qo[nq] = (uint32_t)(ntask - t0);
    free(sord);
    free(qst);
    if (ntask > 0xFFFFFFFFull) {
      adatm_set_error("persistent launch: too many tasks");
      rc = SPLATT_ERROR_BADINPUT;
    }
    if (nq + b->p_nstage[r] > b->ctr_words) b->ctr_words = nq + b->p_nstage[r];
  }
  if (rc == SPLATT_SUCCESS) {
    size_t const ns = all.n ? all.n : 1, nt = ntask ? ntask : 1;
    if (hipMalloc((void **)&b->dstages, ns * sizeof(adatm_kstage)) != hipSuccess ||
        hipMalloc((void **)&b->dtasks, nt * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc((void **)&b->dqoff, (size_t)nr * (nq + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->dxmap, 16 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->dctr, (size_t)(b->ctr_words ? b->ctr_words : 1) * 4) !=
            hipSuccess ||
        hipMalloc((void **)&b->dstat, (size_t)2 * nq * 4) != hipSuccess ||
        (all.n && hipMemcpy(b->dstages, all.st, all.n * sizeof(adatm_kstage),
                            hipMemcpyHostToDevice) != hipSuccess) ||
        (ntask && hipMemcpy(b->dtasks, tasks, ntask * sizeof(uint64_t),
                            hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(b->dqoff, qoff, (size_t)nr * (nq + 1) * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->dxmap, b->xmap, 16 * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess)
      rc = hip_fail(hipGetLastError(), "persistent launch descriptors");
  }
  if (rc != SPLATT_SUCCESS) { /* no half-built persistent mode */
    hipFree(b->dstages);
    hipFree(b->dtasks);
    hipFree(b->dqoff);
    hipFree(b->dxmap);
    hipFree(b->dctr);
    hipFree(b->dstat);
    b->dstages = NULL;
    b->dtasks = NULL;
    b->dqoff = NULL;
    b->dxmap = b->dctr = b->dstat = NULL;
  }
  free(all.st);
  free(all.cost);
  free(tasks);
  free(qoff);
  free(first);
  free(len);
  free(queue);
  b->spin = PERSIST_SPIN;
  return rc;
}

/* ---- grouped form ---------------------------------------------------------
 * Each round's recorded stages run as ordinary launches, one per depth of the
 * chains and kind of stage (type and template arguments): every instance's
 * first stages of one kind in one launch, then the next kind, then the
 * second stages, and so on.  The stream orders a depth after the one before,
 * so no device-side waiting or cache maintenance is needed, and each launch
 * spreads its instances' blocks over the whole device. */

static int same_kind(adatm_kstage const *a, adatm_kstage const *b)
{
  return a->type == b->type && a->leaf == b->leaf && a->K == b->K && a->side == b->side &&
         a->d == b->d && a->zs == b->zs;
}

/* every instance's chain of stages for round r appended to `all`: chain i
   is all->st[first[i] .. first[i] + len[i]), each stage waiting for the one
   before (round-local indices from s0) */
static int batch_record_round(adatm_batch *b, uint32_t r, krec *all, uint32_t s0,
                              uint32_t *first, uint32_t *len)
{
  int rc = SPLATT_SUCCESS;
  for (uint32_t i = 0; rc == SPLATT_SUCCESS && i < (uint32_t)b->n; ++i) {
    adatm_engine *e = b->e[i];
    first[i] = all->n;
    len[i] = 0;
    if (e->N <= r) continue;
    e->rec = all;
    rc = mttkrp_dev(e, r);
    e->rec = NULL;
    len[i] = all->n - first[i];
    for (uint32_t k = first[i]; rc == SPLATT_SUCCESS && k < all->n; ++k) {
      adatm_kstage *x = all->st + k;
      if ((x->type == KS_REDUCE && x->K > 5) || (x->type == KS_SCATTER && x->d > 5)) {
        adatm_set_error("batched stages: a stage deeper than 5 levels");
        rc = SPLATT_ERROR_BADINPUT;
      }
      x->dep = (k == first[i]) ? ADATM_KS_NODEP : k - 1 - s0;
    }
  }
  return rc;
}

static int batch_grouped_build(adatm_batch *b)
{
  uint32_t const nr = b->nround, n = (uint32_t)b->n;
  krec all = {0};
  uint64_t *bmap = NULL, nbm = 0, bm_cap = 0;
  uint32_t ngl = 0, gl_cap = 0;
  uint32_t *first = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *len = (uint32_t *)malloc(n * sizeof(uint32_t));
  free(b->gl); /* a failed earlier attempt */
  free(b->gl0);
  b->gl = NULL;
  b->gl0 = (uint32_t *)calloc(nr + 1, sizeof(uint32_t));
  int rc = (first && len && b->gl0) ? SPLATT_SUCCESS : SPLATT_ERROR_NOMEMORY;
  for (uint32_t r = 0; rc == SPLATT_SUCCESS && r < nr; ++r) {
    uint32_t const s0 = all.n;
    b->gl0[r] = ngl;
    rc = batch_record_round(b, r, &all, s0, first, len);
    uint32_t maxlen = 0;
    for (uint32_t i = 0; rc == SPLATT_SUCCESS && i < n; ++i)
      if (len[i] > maxlen) maxlen = len[i];
    for (uint32_t d = 0; rc == SPLATT_SUCCESS && d < maxlen; ++d) {
      /* the kinds of depth d in the order they first occur; one launch each */
      for (uint32_t i0 = 0; rc == SPLATT_SUCCESS && i0 < n; ++i0) {
        if (d >= len[i0]) continue;
        adatm_kstage const *k0 = all.st + first[i0] + d;
        int seen = 0;
        for (uint32_t i = 0; i < i0 && !seen; ++i)
          seen = d < len[i] && same_kind(all.st + first[i] + d, k0);
        if (seen) continue;
        if (ngl == gl_cap) {
          uint32_t const cap = gl_cap ? 2 * gl_cap : 64;
          glaunch *ng = (glaunch *)realloc(b->gl, cap * sizeof(glaunch));
          if (!ng) {
            rc = SPLATT_ERROR_NOMEMORY;
            continue;
          }
          b->gl = ng;
          gl_cap = cap;
        }
        glaunch *L = b->gl + ngl++;
        L->kind = *k0;
        L->bmap0 = nbm;
        L->nb = 0;
        for (uint32_t i = i0; rc == SPLATT_SUCCESS && i < n; ++i) {
          if (d >= len[i] || !same_kind(all.st + first[i] + d, k0)) continue;
          uint32_t const si = first[i] + d, nbk = all.st[si].nblocks;
          if (nbm + nbk > bm_cap) {
            uint64_t cap = bm_cap ? 2 * bm_cap : 4096;
            while (cap < nbm + nbk) cap *= 2;
            uint64_t *nb = (uint64_t *)realloc(bmap, cap * sizeof(uint64_t));
            if (!nb) {
              rc = SPLATT_ERROR_NOMEMORY;
              continue;
            }
            bmap = nb;
            bm_cap = cap;
          }
          for (uint32_t v = 0; v < nbk; ++v) bmap[nbm++] = ((uint64_t)si << 32) | v;
          L->nb += nbk;
        }
      }
    }
  }
  if (rc == SPLATT_SUCCESS) {
    b->gl0[nr] = ngl;
    size_t const ns = all.n ? all.n : 1, nm = nbm ? nbm : 1;
    if (hipMalloc((void **)&b->gst, ns * sizeof(adatm_kstage)) != hipSuccess ||
        hipMalloc((void **)&b->gbmap, nm * sizeof(uint64_t)) != hipSuccess ||
        (all.n && hipMemcpy(b->gst, all.st, all.n * sizeof(adatm_kstage),
                            hipMemcpyHostToDevice) != hipSuccess) ||
        (nbm && hipMemcpy(b->gbmap, bmap, nbm * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                    hipSuccess))
      rc = hip_fail(hipGetLastError(), "grouped launch descriptors");
  }
  if (rc != SPLATT_SUCCESS) {
    hipFree(b->gst);
    hipFree(b->gbmap);
    b->gst = NULL;
    b->gbmap = NULL;
  }
  free(all.st);
  free(all.cost);
  free(bmap);
  free(first);
  free(len);
  return rc;
}

static int batch_enqueue_grouped_form(adatm_batch *b)
{
  for (uint32_t r = 0; r < b->nround; ++r) {
    batch_round const *B = b->round + r;
    for (uint32_t l = b->gl0[r]; l < b->gl0[r + 1]; ++l) {
      glaunch const *L = b->gl + l;
      KCHK(adatm_k_grouped(b->s, b->e[0]->prec, &L->kind, b->gst, b->gbmap + L->bmap0, L->nb),
           "grouped MTTKRP launch");
    }
    KCHK(adatm_k_dense_grouped(b->s, b->e[0]->prec, b->ddesc + B->desc0, B->n, B->Rmax,
                                b->dmap + B->mm_off, B->nb_mm, b->dmap + B->gr_off,
                                B->nb_gr, b->dmap + B->sp_off, B->nb_sp),
          "grouped update");
  }
  KCHK(adatm_k_fit_grouped(b->s, b->e[0]->prec, b->ddesc + b->fit0, (uint32_t)b->n,
                           b->dmap + b->in_off, b->nb_in),
       "grouped fit");
  return SPLATT_SUCCESS;
}

/* the persistent sweep: per round, the counters cleared, one persistent
 * launch of every instance's MTTKRP, one grouped dense update; one grouped
 * fit last */
static int batch_enqueue_persist(adatm_batch *b)
{
  uint32_t const nq = b->nq;
  for (uint32_t r = 0; r < b->nround; ++r) {
    batch_round const *B = b->round + r;
    KCHK(adatm_k_zero(b->s, b->dctr, (size_t)(nq + b->p_nstage[r]) * 4), "persistent counters");
    KCHK(adatm_k_persist(b->s, b->e[0]->prec, b->pgrid, b->dstages + b->p_stage0[r],
                         b->dtasks + b->p_task0[r], b->dqoff + (size_t)r * (nq + 1), b->dxmap,
                         nq, b->dctr, b->dstat, b->spin),
         "persistent launch");
    KCHK(adatm_k_dense_grouped(b->s, b->e[0]->prec, b->ddesc + B->desc0, B->n, B->Rmax,
                                b->dmap + B->mm_off, B->nb_mm, b->dmap + B->gr_off,
                                B->nb_gr, b->dmap + B->sp_off, B->nb_sp),
          "grouped update");
  }
  KCHK(adatm_k_fit_grouped(b->s, b->e[0]->prec, b->ddesc + b->fit0, (uint32_t)b->n,
                           b->dmap + b->in_off, b->nb_in),
       "grouped fit");
  return SPLATT_SUCCESS;
}

/* the grouped sweep: per round, the instances' MTTKRPs forked on their
 * branch streams, then one grouped dense update on the batch stream; one
 * grouped fit last */
static int batch_enqueue_grouped(adatm_batch *b)
{
  for (uint32_t r = 0; r < b->nround; ++r) {
    batch_round const *B = b->round + r;
    if (hipEventRecord(b->fork, b->s) != hipSuccess) return SPLATT_ERROR_DEVICE;
    for (int i = 0; i < b->n; ++i) {
      if (b->e[i]->N <= r) continue;
      int const rc = batch_branch(b, i, (int)r);
      if (rc != SPLATT_SUCCESS) return rc;
    }
    KCHK(adatm_k_dense_grouped(b->s, b->e[0]->prec, b->ddesc + B->desc0, B->n, B->Rmax,
                                b->dmap + B->mm_off, B->nb_mm, b->dmap + B->gr_off,
                                B->nb_gr, b->dmap + B->sp_off, B->nb_sp),
          "grouped update");
  }
  KCHK(adatm_k_fit_grouped(b->s, b->e[0]->prec, b->ddesc + b->fit0, (uint32_t)b->n,
                           b->dmap + b->in_off, b->nb_in),
       "grouped fit");
  return SPLATT_SUCCESS;
}

int adatm_batch_create(adatm_engine *const *engines, int n, adatm_batch **out)
{
  *out = NULL;
  if (!engines || n < 1) return SPLATT_ERROR_BADINPUT;
  for (int i = 0; i < n; ++i)
    if (!engines[i] || engines[i]->device != engines[0]->device) {
      adatm_set_error("batch: engine %d is missing or on another device", i);
      return SPLATT_ERROR_BADINPUT;
    }
  adatm_batch *b = (adatm_batch *)calloc(1, sizeof(adatm_batch));
  if (!b) return SPLATT_ERROR_NOMEMORY;
  b->n = n;
  b->device = engines[0]->device;
  b->e = (adatm_engine **)malloc((size_t)n * sizeof(adatm_engine *));
  b->join = (hipEvent_t *)calloc((size_t)n, sizeof(hipEvent_t));
  b->bs = (hipStream_t *)calloc((size_t)n, sizeof(hipStream_t));
  if (!b->e || !b->join || !b->bs) {
    adatm_batch_destroy(b);
    return SPLATT_ERROR_NOMEMORY;
  }
  memcpy(b->e, engines, (size_t)n * sizeof(adatm_engine *));
  /* grouped dense updates need one precision across the instances */
  b->grouped = 1;
  for (int i = 1; i < n; ++i)
    if (engines[i]->prec != engines[0]->prec) b->grouped = 0;
  /* grouped launches where they apply (measured fastest), else the branches */
  b->want_form = ADATM_FORM_GROUPED;
  int rc = SPLATT_SUCCESS;
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    rc = hip_fail(hipGetLastError(), "batch stream");
  for (int i = 0; rc == SPLATT_SUCCESS && i < n; ++i)
    if (hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&b->bs[i], hipStreamNonBlocking) != hipSuccess)
      rc = hip_fail(hipGetLastError(), "batch event / branch stream");
  if (rc != SPLATT_SUCCESS) {
    adatm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return SPLATT_SUCCESS;
}

static int batch_graph(adatm_batch *b)
{
  if (b->gvalid) return SPLATT_SUCCESS;
  if (b->grouped && !b->ddesc) {
    int const rc = batch_rounds(b);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  int const ok = batch_persist_ok(b);
  if (b->want_form == ADATM_FORM_PERSISTENT && !b->dstages && ok) {
    int const rc = batch_persist_build(b);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  if (b->want_form == ADATM_FORM_GROUPED && !b->gst && ok) {
    int const rc = batch_grouped_build(b);
    if (rc != SPLATT_SUCCESS) return rc;
  }
  b->form = (b->want_form == ADATM_FORM_PERSISTENT && b->dstages)  ? ADATM_FORM_PERSISTENT
            : (b->want_form == ADATM_FORM_GROUPED && b->gst) ? ADATM_FORM_GROUPED
                                                              : ADATM_FORM_FORKED;
  HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal), "begin capture");
  int rc = SPLATT_SUCCESS;
  if (b->form == ADATM_FORM_PERSISTENT)
    rc = batch_enqueue_persist(b);
  else if (b->form == ADATM_FORM_GROUPED)
    rc = batch_enqueue_grouped_form(b);
  else if (b->grouped)
    rc = batch_enqueue_grouped(b);
  else if (hipEventRecord(b->fork, b->s) != hipSuccess)
    rc = SPLATT_ERROR_DEVICE;
  for (int i = 0; !b->grouped && rc == SPLATT_SUCCESS && i < b->n; ++i)
    rc = batch_branch(b, i, -1);
  hipGraph_t g = NULL;
  hipError_t const ec = hipStreamEndCapture(b->s, &g);
  if (rc != SPLATT_SUCCESS || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    if (rc == SPLATT_SUCCESS) rc = hip_fail(ec, "end capture");
    else if (!adatm_last_error()[0]) adatm_set_error("batch capture failed");
    return rc;
  }
  hipError_t const ei = hipGraphInstantiate(&b->gexec, g, NULL, NULL, 0);
  if (ei != hipSuccess) {
    hipGraphDestroy(g);
    return hip_fail(ei, "batch graph instantiate");
  }
  b->graph = g;
  b->gvalid = 1;
  return SPLATT_SUCCESS;
}

/* niters batched sweeps from every instance's current factors: one graph
 * launch per sweep.  *ms_total = device time of the sweeps (events on the
 * batch stream); fits (nullable, n x niters row-major) = every instance's
 * fit after every sweep. */
int adatm_batch_cpd(adatm_batch *b, uint32_t niters, double *ms_total, double *fits)
{
  if (!b || niters == 0) return SPLATT_ERROR_BADINPUT;
  HIPCHK(hipSetDevice(b->device), "hipSetDevice");
  for (int i = 0; i < b->n; ++i) {
    if (niters > b->e[i]->fits_cap) return SPLATT_ERROR_BADINPUT;
    int const rc = cpd_prepare(b->e[i]);
    if (rc != SPLATT_SUCCESS) return rc;
    HIPCHK(hipStreamSynchronize(b->e[i]->stream), "prepare sync");
  }
  int rc = batch_graph(b);
  if (rc != SPLATT_SUCCESS) return rc;
  if (b->form == ADATM_FORM_PERSISTENT)
    HIPCHK(hipMemset(b->dstat, 0, (size_t)2 * b->nq * 4), "persistent statistics");
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0), "event");
  hipError_t const et1 = hipEventCreate(&t1);
  if (et1 != hipSuccess) {
    hipEventDestroy(t0);
    return hip_fail(et1, "event");
  }
  hipEventRecord(t0, b->s);
  for (uint32_t it = 0; it < niters; ++it)
    if (hipGraphLaunch(b->gexec, b->s) != hipSuccess) {
      rc = hip_fail(hipGetLastError(), "batch graph launch");
      break;
    }
  hipEventRecord(t1, b->s);
  hipError_t const es = hipEventSynchronize(t1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (rc != SPLATT_SUCCESS) return rc;
  HIPCHK(es, "batch sync");
  for (int i = 0; i < b->n; ++i) b->e[i]->nsweep = niters; /* the device counters */
  if (b->form == ADATM_FORM_PERSISTENT) {
    /* every queue served all its tasks, without a wait that gave up */
    uint32_t st[32];
    HIPCHK(hipMemcpy(st, b->dstat, (size_t)2 * b->nq * 4, hipMemcpyDeviceToHost),
           "persistent statistics");
    for (uint32_t q = 0; q < b->nq; ++q) {
      if (st[q]) {
        adatm_set_error("persistent launch: a dependency wait of queue %u gave up", q);
        return SPLATT_ERROR_DEVICE;
      }
      if ((uint64_t)st[b->nq + q] != b->p_served[q] * niters) {
        adatm_set_error("persistent launch: queue %u ran %u of %llu tasks", q,
                        st[b->nq + q], (unsigned long long)(b->p_served[q] * niters));
        return SPLATT_ERROR_DEVICE;
      }
    }
  }
  if (ms_total) *ms_total = ms;
  if (fits)
    for (int i = 0; i < b->n; ++i)
      HIPCHK(hipMemcpy(fits + (size_t)i * niters, b->e[i]->fits, niters * sizeof(double),
                       hipMemcpyDeviceToHost),
             "batch fits");
  int worst = 0;
  for (int i = 0; i < b->n; ++i) {
    int st = 0;
    rc = read_status(b->e[i], &st);
    if (rc != SPLATT_SUCCESS) return rc;
    if (st > worst) worst = st;
  }
  return status_rc(worst);
}

int adatm_batch_set_form(adatm_batch *b, int form)
{
  if (!b || form < ADATM_FORM_FORKED || form > ADATM_FORM_GROUPED) return SPLATT_ERROR_BADINPUT;
  if (form != ADATM_FORM_FORKED && !batch_persist_ok(b)) {
    adatm_set_error("batch: the persistent and grouped forms need one precision, "
                    "8 <= R <= 32 and at most 6 modes per instance");
    return SPLATT_ERROR_BADINPUT;
  }
  if (form != b->want_form && b->gvalid) {
    hipGraphExecDestroy(b->gexec);
    hipGraphDestroy(b->graph);
    b->gvalid = 0;
  }
  b->want_form = form;
  return SPLATT_SUCCESS;
}

int adatm_batch_form(adatm_batch const *b) { return b ? b->form : ADATM_FORM_FORKED; }

int adatm_batch_set_persistent(adatm_batch *b, int on)
{
  return adatm_batch_set_form(b, on ? ADATM_FORM_PERSISTENT : ADATM_FORM_FORKED);
}

int adatm_batch_persistent(adatm_batch const *b)
{
  return b && b->form == ADATM_FORM_PERSISTENT;
}

int adatm_batch_persist_info(adatm_batch const *b, uint32_t *nq, uint32_t *xmap,
                             uint32_t *grid)
{
  if (!b) return SPLATT_ERROR_BADINPUT;
  if (nq) *nq = b->nq;
  if (xmap) memcpy(xmap, b->xmap, sizeof(b->xmap));
  if (grid) *grid = b->pgrid;
  return SPLATT_SUCCESS;
}

/* The unbatched comparison: each instance's niters sweeps (its own graph)
 * one instance after another, synchronised between instances. */
int adatm_batch_sequential(adatm_batch *b, uint32_t niters, double *ms_total)
{
  if (!b || niters == 0) return SPLATT_ERROR_BADINPUT;
  double total = 0.0;
  for (int i = 0; i < b->n; ++i) {
    double ms = 0.0;
    int const rc = adatm_engine_run_timed(b->e[i], 1, niters, 1, &ms);
    if (rc != SPLATT_SUCCESS) return rc;
    total += ms;
  }
  *ms_total = total;
  return SPLATT_SUCCESS;
}
