/*
 * Host construction helpers: stable LSD radix sort of coordinates,
 * canonicalisation (duplicate merge, empty-index compaction), fiber counts,
 * CSF structure and the segmented-reduction work decomposition.
 *
 * Reference behaviour restated here:
 *   - tt_sort (SPLATT, called at adatm_rcsf.c:179, adatm_bench.c:179/222):
 *     lexicographic order of the coordinates by a mode permutation;
 *   - count_nfibs (adatm_rcsf.c:410-500): F_l = number of distinct
 *     length-(l+1) prefixes of the sorted coordinates;
 *   - p_mk_outerptr / p_mk_fptr / p_csf_alloc_untiled (adatm_rcsf.c:21-209):
 *     one CSF node per distinct prefix, fptr = first child, fids = the
 *     coordinate of the level's mode, vals/leaf ids copied in sorted order;
 *   - the untiled CSF drops root ids (adatm_rcsf.c:51-56) so the root output
 *     row is the slice ordinal; compaction of empty indices makes the slice
 *     ordinal equal the root coordinate.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "adatm_host.h"

static _Thread_local char g_err[512];

void adatm_set_error(char const *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

char const *adatm_last_error(void) { return g_err; }

double adatm_wtime(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int nthreads_max(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static uint32_t bits_for(uint64_t dim)
{
  uint32_t b = 0;
  while (b < 32 && (1ull << b) < dim) ++b;
  return b;
}

#define RADIX_BITS 11
#define RADIX (1u << RADIX_BITS)

/* One stable counting-sort pass of (key, perm) by digit `shift`. */
static void radix_pass(uint64_t n, uint32_t shift, uint32_t const *kin,
                       uint32_t const *pin, uint32_t *kout, uint32_t *pout,
                       uint64_t *hist, int nt)
{
  uint64_t const chunk = (n + (uint64_t)nt - 1) / (uint64_t)nt;
  #pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int const t = omp_get_thread_num();
#else
    int const t = 0;
#endif
    uint64_t *h = hist + (uint64_t)t * RADIX;
    memset(h, 0, RADIX * sizeof(uint64_t));
    uint64_t const lo = (uint64_t)t * chunk;
    uint64_t const hi = lo + chunk < n ? lo + chunk : n;
    for (uint64_t j = lo; j < hi; ++j) h[(kin[j] >> shift) & (RADIX - 1)]++;
    #pragma omp barrier
    #pragma omp single
    {
      uint64_t run = 0;
      for (uint32_t b = 0; b < RADIX; ++b)
        for (int tt = 0; tt < nt; ++tt) {
          uint64_t const c = hist[(uint64_t)tt * RADIX + b];
          hist[(uint64_t)tt * RADIX + b] = run;
          run += c;
        }
    }
    for (uint64_t j = lo; j < hi; ++j) {
      uint64_t const dst = h[(kin[j] >> shift) & (RADIX - 1)]++;
      kout[dst] = kin[j];
      pout[dst] = pin[j];
    }
  }
}

int adatm_lex_sort_perm(adatm_coo const *t, uint32_t const *order,
                        uint32_t nkeys, uint32_t *perm)
{
  uint64_t const n = t->nnz;
  int const nt = nthreads_max();
  uint32_t *k0 = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *k1 = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *p1 = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t *hist = (uint64_t *)malloc((size_t)nt * RADIX * sizeof(uint64_t));
  if (!k0 || !k1 || !p1 || !hist) {
    free(k0); free(k1); free(p1); free(hist);
    adatm_set_error("host sort: out of memory (nnz=%llu)",
                    (unsigned long long)n);
    return SPLATT_ERROR_NOMEMORY;
  }
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j) perm[j] = (uint32_t)j;

  for (int64_t kk = (int64_t)nkeys - 1; kk >= 0; --kk) {
    uint32_t const *key = t->ind[order[kk]];
    uint32_t const nb = bits_for(t->dims[order[kk]]);
    if (nb == 0) continue; /* a mode of length one never reorders */
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j) k0[j] = key[perm[j]];
    uint32_t *ka = k0, *kb = k1, *pa = perm, *pb = p1;
    for (uint32_t shift = 0; shift < nb; shift += RADIX_BITS) {
      radix_pass(n, shift, ka, pa, kb, pb, hist, nt);
      uint32_t *tmp = ka; ka = kb; kb = tmp;
      tmp = pa; pa = pb; pb = tmp;
    }
    if (pa != perm) memcpy(perm, pa, n * sizeof(uint32_t));
  }
  free(k0); free(k1); free(p1); free(hist);
  return SPLATT_SUCCESS;
}

/* Exclusive prefix sum of flags (0/1 bytes) in parallel; returns total. */
static uint64_t flag_scan(uint8_t const *flag, uint64_t n, uint32_t *pos)
{
  int const nt = nthreads_max();
  uint64_t *part = (uint64_t *)calloc((size_t)nt + 1, sizeof(uint64_t));
  uint64_t const chunk = (n + (uint64_t)nt - 1) / (uint64_t)nt;
  #pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int const t = omp_get_thread_num();
#else
    int const t = 0;
#endif
    uint64_t const lo = (uint64_t)t * chunk;
    uint64_t const hi = lo + chunk < n ? lo + chunk : n;
    uint64_t c = 0;
    for (uint64_t j = lo; j < hi; ++j) c += flag[j];
    part[t + 1] = c;
    #pragma omp barrier
    #pragma omp single
    for (int tt = 0; tt < nt; ++tt) part[tt + 1] += part[tt];
    uint64_t run = part[t];
    for (uint64_t j = lo; j < hi; ++j) {
      pos[j] = (uint32_t)run;
      run += flag[j];
    }
  }
  uint64_t const total = part[nt];
  free(part);
  return total;
}

int adatm_coo_canonicalize(adatm_coo *t, uint32_t **indmap_out)
{
  uint32_t const N = t->nmodes;
  uint64_t const n = t->nnz;
  if (N < 1 || N > ADATM_MAX_NMODES || n == 0 || n >= (1ull << 32)) {
    adatm_set_error("canonicalize: bad tensor (N=%u nnz=%llu)", N,
                    (unsigned long long)n);
    return SPLATT_ERROR_BADINPUT;
  }
  for (uint32_t m = 0; m < N; ++m) {
    uint32_t const *ix = t->ind[m];
    uint64_t const dim = t->dims[m];
    int bad = 0;
    #pragma omp parallel for reduction(| : bad)
    for (int64_t j = 0; j < (int64_t)n; ++j) bad |= (ix[j] >= dim);
    if (bad) {
      adatm_set_error("canonicalize: index out of range in mode %u", m);
      return SPLATT_ERROR_BADINPUT;
    }
  }
  uint32_t order[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < N; ++m) order[m] = m;
  uint32_t *perm = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint8_t *start = (uint8_t *)malloc(n);
  uint32_t *pos = (uint32_t *)malloc(n * sizeof(uint32_t));
  if (!perm || !start || !pos) {
    free(perm); free(start); free(pos);
    return SPLATT_ERROR_NOMEMORY;
  }
  int rc = adatm_lex_sort_perm(t, order, N, perm);
  if (rc != SPLATT_SUCCESS) {
    free(perm); free(start); free(pos);
    return rc;
  }
  /* run starts in sorted order */
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j) {
    uint8_t s = (j == 0);
    if (!s)
      for (uint32_t m = 0; m < N; ++m)
        if (t->ind[m][perm[j]] != t->ind[m][perm[j - 1]]) {
          s = 1;
          break;
        }
    start[j] = s;
  }
  uint64_t const nu = flag_scan(start, n, pos);
  /* run boundaries */
  uint32_t *rb = (uint32_t *)malloc((nu + 1) * sizeof(uint32_t));
  double *nv = (double *)malloc(nu * sizeof(double));
  if (!rb || !nv) {
    free(perm); free(start); free(pos); free(rb); free(nv);
    return SPLATT_ERROR_NOMEMORY;
  }
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j)
    if (start[j]) rb[pos[j]] = (uint32_t)j;
  rb[nu] = (uint32_t)n;
  /* duplicates summed in draw order: a stable sort keeps draw order */
  #pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < (int64_t)nu; ++r) {
    double s = 0.0;
    for (uint32_t j = rb[r]; j < rb[r + 1]; ++j) s += t->vals[perm[j]];
    nv[r] = s;
  }
  for (uint32_t m = 0; m < N; ++m) {
    uint32_t *ni = (uint32_t *)malloc(nu * sizeof(uint32_t));
    if (!ni) {
      free(perm); free(start); free(pos); free(rb); free(nv);
      return SPLATT_ERROR_NOMEMORY;
    }
    uint32_t const *oi = t->ind[m];
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)nu; ++r) ni[r] = oi[perm[rb[r]]];
    free(t->ind[m]);
    t->ind[m] = ni;
  }
  free(t->vals);
  t->vals = nv;
  t->nnz = nu;
  free(perm); free(start); free(pos); free(rb);

  /* drop empty indices (monotone relabel keeps the sort order) */
  for (uint32_t m = 0; m < N; ++m) {
    uint64_t const dim = t->dims[m];
    uint8_t *occ = (uint8_t *)calloc(dim, 1);
    uint32_t *nid = (uint32_t *)malloc(dim * sizeof(uint32_t));
    if (!occ || !nid) {
      free(occ); free(nid);
      return SPLATT_ERROR_NOMEMORY;
    }
    uint32_t *ix = t->ind[m];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)nu; ++j) occ[ix[j]] = 1;
    uint64_t const nd = flag_scan(occ, dim, nid);
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)nu; ++j) ix[j] = nid[ix[j]];
    if (indmap_out) {
      uint32_t *map = (uint32_t *)malloc((nd ? nd : 1) * sizeof(uint32_t));
      if (map)
        for (uint64_t i = 0; i < dim; ++i)
          if (occ[i]) map[nid[i]] = (uint32_t)i;
      indmap_out[m] = map;
    }
    t->dims[m] = nd;
    free(occ);
    free(nid);
  }
  return SPLATT_SUCCESS;
}

/* first differing level (in `order`) between sorted positions j-1 and j */
static void first_diff(adatm_coo const *t, uint32_t const *order,
                       uint32_t const *perm, uint8_t *fd)
{
  uint32_t const N = t->nmodes;
  uint64_t const n = t->nnz;
  fd[0] = 0;
  #pragma omp parallel for schedule(static)
  for (int64_t j = 1; j < (int64_t)n; ++j) {
    uint8_t l = (uint8_t)N;
    for (uint32_t k = 0; k < N; ++k)
      if (t->ind[order[k]][perm[j]] != t->ind[order[k]][perm[j - 1]]) {
        l = (uint8_t)k;
        break;
      }
    fd[j] = l;
  }
}

int adatm_count_fibers_host(adatm_coo const *t, uint32_t const *order,
                            uint64_t *nfibs)
{
  uint32_t const N = t->nmodes;
  uint64_t const n = t->nnz;
  uint32_t *perm = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint8_t *fd = (uint8_t *)malloc(n);
  if (!perm || !fd) {
    free(perm); free(fd);
    return SPLATT_ERROR_NOMEMORY;
  }
  int rc = adatm_lex_sort_perm(t, order, N, perm);
  if (rc != SPLATT_SUCCESS) {
    free(perm); free(fd);
    return rc;
  }
  first_diff(t, order, perm, fd);
  uint64_t hist[ADATM_MAX_NMODES + 1] = {0};
  int const nt = nthreads_max();
  uint64_t *th = (uint64_t *)calloc((size_t)nt * (ADATM_MAX_NMODES + 1),
                                    sizeof(uint64_t));
  #pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    uint64_t *h = th + (uint64_t)omp_get_thread_num() * (ADATM_MAX_NMODES + 1);
#else
    uint64_t *h = th;
#endif
    #pragma omp for schedule(static)
    for (int64_t j = 1; j < (int64_t)n; ++j) h[fd[j]]++;
  }
  for (int tt = 0; tt < nt; ++tt)
    for (uint32_t l = 0; l <= ADATM_MAX_NMODES; ++l)
      hist[l] += th[(uint64_t)tt * (ADATM_MAX_NMODES + 1) + l];
  free(th);
  uint64_t run = 1;
  for (uint32_t l = 0; l + 1 < N; ++l) {
    run += hist[l];
    nfibs[l] = run;
  }
  nfibs[N - 1] = n;
  free(perm);
  free(fd);
  return SPLATT_SUCCESS;
}

void adatm_hcsf_free(adatm_hcsf *c)
{
  if (!c) return;
  for (uint32_t l = 0; l < ADATM_MAX_NMODES; ++l) {
    free(c->fptr[l]);
    free(c->fids[l]);
    free(c->par[l]);
    c->fptr[l] = c->fids[l] = c->par[l] = NULL;
  }
  free(c->vals);
  c->vals = NULL;
}

int adatm_hcsf_build(adatm_coo const *t, uint32_t const *order,
                     adatm_hcsf *out)
{
  uint32_t const N = t->nmodes;
  uint64_t const n = t->nnz;
  memset(out, 0, sizeof(*out));
  out->N = N;
  out->nnz = n;
  memcpy(out->order, order, N * sizeof(uint32_t));
  uint32_t *perm = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint8_t *fd = (uint8_t *)malloc(n);
  uint8_t *flag = (uint8_t *)malloc(n);
  uint32_t *pos = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *S[ADATM_MAX_NMODES] = {0}; /* node start positions per level */
  int rc = SPLATT_ERROR_NOMEMORY;
  if (!perm || !fd || !flag || !pos) goto done;
  rc = adatm_lex_sort_perm(t, order, N, perm);
  if (rc != SPLATT_SUCCESS) goto done;
  rc = SPLATT_ERROR_NOMEMORY;
  first_diff(t, order, perm, fd);

  out->vals = (double *)malloc(n * sizeof(double));
  out->fids[N - 1] = (uint32_t *)malloc(n * sizeof(uint32_t));
  if (!out->vals || !out->fids[N - 1]) goto done;
  {
    uint32_t const *leaf = t->ind[order[N - 1]];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
      out->vals[j] = t->vals[perm[j]];
      out->fids[N - 1][j] = leaf[perm[j]];
    }
  }
  out->F[N - 1] = n;
  for (uint32_t l = 0; l + 1 < N; ++l) {
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j)
      flag[j] = (uint8_t)(j == 0 || fd[j] <= l);
    uint64_t const F = flag_scan(flag, n, pos);
    out->F[l] = F;
    S[l] = (uint32_t *)malloc(F * sizeof(uint32_t));
    out->fids[l] = (uint32_t *)malloc(F * sizeof(uint32_t));
    if (!S[l] || !out->fids[l]) goto done;
    uint32_t const *ix = t->ind[order[l]];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j)
      if (flag[j]) {
        S[l][pos[j]] = (uint32_t)j;
        out->fids[l][pos[j]] = ix[perm[j]];
      }
  }
  /* fptr[l][p] = first level-(l+1) node under p; par[l+1][c] = parent */
  for (uint32_t l = 0; l + 1 < N; ++l) {
    uint64_t const Fp = out->F[l], Fc = out->F[l + 1];
    out->fptr[l] = (uint32_t *)malloc((Fp + 1) * sizeof(uint32_t));
    out->par[l + 1] = (uint32_t *)malloc(Fc * sizeof(uint32_t));
    if (!out->fptr[l] || !out->par[l + 1]) goto done;
    uint32_t *fp = out->fptr[l], *pa = out->par[l + 1];
    uint64_t c = 0;
    for (uint64_t p = 0; p < Fp; ++p) {
      uint32_t const sp = S[l][p];
      uint32_t const ep = (p + 1 < Fp) ? S[l][p + 1] : (uint32_t)n;
      if (l + 1 == N - 1) {
        fp[p] = sp;
        for (uint32_t j = sp; j < ep; ++j) pa[j] = (uint32_t)p;
        c = ep;
      } else {
        fp[p] = (uint32_t)c;
        while (c < Fc && S[l + 1][c] < ep) pa[c++] = (uint32_t)p;
      }
    }
    fp[Fp] = (uint32_t)Fc;
  }
  rc = SPLATT_SUCCESS;
done:
  for (uint32_t l = 0; l < ADATM_MAX_NMODES; ++l) free(S[l]);
  free(perm);
  free(fd);
  free(flag);
  free(pos);
  if (rc != SPLATT_SUCCESS) {
    adatm_hcsf_free(out);
    adatm_set_error("host CSF build failed (%d)", rc);
  }
  return rc;
}


/* Fiber counts on the device (radix sort + boundary scan). */
int adatm_count_fibers_device(adatm_coo const *t, uint32_t const *order,
                              uint64_t *nfibs, int device)
{
  return adatm_count_fibers_device_many(t, order, 1, nfibs, device);
}

/* Allocate an empty coordinate tensor (malloc'd, freed by adatm_coo_free). */
int adatm_coo_alloc(uint32_t nmodes, uint64_t nnz, uint64_t const *dims,
                    adatm_coo *out)
{
  memset(out, 0, sizeof(*out));
  if (nmodes < 1 || nmodes > ADATM_MAX_NMODES || nnz >= (1ull << 32))
    return SPLATT_ERROR_BADINPUT;
  out->nmodes = nmodes;
  out->nnz = nnz;
  for (uint32_t m = 0; m < nmodes; ++m) {
    out->dims[m] = dims[m];
    out->ind[m] = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    if (!out->ind[m]) {
      adatm_coo_free(out);
      return SPLATT_ERROR_NOMEMORY;
    }
  }
  out->vals = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
  if (!out->vals) {
    adatm_coo_free(out);
    return SPLATT_ERROR_NOMEMORY;
  }
  return SPLATT_SUCCESS;
}

/* Sum of squares in a fixed association (1024 contiguous blocks summed in
 * order), so the value does not depend on the thread count or timing. */
double adatm_sumsq(double const *v, uint64_t n)
{
  enum { NB = 1024 };
  double part[NB];
  uint64_t const chunk = (n + NB - 1) / NB;
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < NB; ++b) {
    uint64_t const lo = (uint64_t)b * chunk;
    uint64_t const hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (uint64_t j = lo; j < hi; ++j) s += v[j] * v[j];
    part[b] = s;
  }
  double s = 0.0;
  for (int b = 0; b < NB; ++b) s += part[b];
  return s;
}
