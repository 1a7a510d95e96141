/*
 * Synthetic workload generators and factor initialisation.
 *
 * Counter-based (Philox4x32-10, Salmon et al., "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11), so draw j of a tensor depends only on (seed, j)
 * and the result is identical for any thread count.  The reference reads
 * tensors from files (cmd_bench.c:234, tt_read) and draws factors with
 * SPLATT's mat_rand (cmd_bench.c:261, adatm_cpd.c:228); both are replaced by
 * these seeded streams so that the CPU and GPU paths see identical inputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "adatm_host.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void adatm_philox4x32_10(uint32_t const ctr_in[4], uint32_t const key_in[2],
                         uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t const p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t const p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t const n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t const n1 = (uint32_t)p1;
    uint32_t const n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t const n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Uniform double in [0,1) with 53 random bits from two 32-bit words. */
static inline double u53(uint32_t a, uint32_t b)
{
  uint64_t const hi = a >> 5, lo = b >> 6;
  return (double)(hi * 67108864ull + lo) * (1.0 / 9007199254740992.0);
}

/* A stream of uniforms for one (seed, counter, tag) triple. */
typedef struct {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  int have; /* doubles left in buf: 2, 1 or 0 */
} ustream;

static inline void us_init(ustream *s, uint64_t seed, uint64_t counter,
                           uint32_t tag)
{
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  s->ctr[0] = (uint32_t)counter;
  s->ctr[1] = (uint32_t)(counter >> 32);
  s->ctr[2] = 0;
  s->ctr[3] = tag;
  s->have = 0;
}

static inline double us_next(ustream *s)
{
  if (s->have == 0) {
    adatm_philox4x32_10(s->ctr, s->key, s->buf);
    s->ctr[2]++;
    s->have = 2;
  }
  double const u = (s->have == 2) ? u53(s->buf[0], s->buf[1])
                                  : u53(s->buf[2], s->buf[3]);
  s->have--;
  return u;
}

static double const TWO_PI = 6.283185307179586476925286766559;

/* Standard normal by Box-Muller; consumes two uniforms, returns one value. */
static inline double us_normal(ustream *s)
{
  double const u1 = us_next(s);
  double const u2 = us_next(s);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(TWO_PI * u2);
}

enum { TAG_COORD = 1, TAG_VALUE = 2, TAG_CENTRE = 3, TAG_FACTOR = 16 };

/* Inverse-CDF table for P(k) ~ (k+1)^-s, k in [0, dim). */
static double *zipf_table(uint64_t dim, double s)
{
  double *cdf = (double *)malloc(dim * sizeof(double));
  if (!cdf) return NULL;
  double acc = 0.0;
  for (uint64_t k = 0; k < dim; ++k) {
    acc += pow((double)(k + 1), -s);
    cdf[k] = acc;
  }
  double const inv = 1.0 / acc;
  for (uint64_t k = 0; k < dim; ++k) cdf[k] *= inv;
  cdf[dim - 1] = 1.0;
  return cdf;
}

/* smallest k with cdf[k] > u */
static inline uint32_t zipf_draw(double const *cdf, uint64_t dim, double u)
{
  uint64_t lo = 0, hi = dim - 1;
  while (lo < hi) {
    uint64_t const mid = lo + (hi - lo) / 2;
    if (cdf[mid] > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  return (uint32_t)lo;
}

int adatm_synth_named(int config, uint64_t ndraws_override,
                      adatm_synth_spec *spec)
{
  memset(spec, 0, sizeof(*spec));
  spec->seed = 0xADA7A000ull + (uint64_t)config;
  switch (config) {
  case 1: {
    uint64_t const d[4] = {2000, 1500, 1000, 500};
    spec->nmodes = 4;
    memcpy(spec->dims, d, sizeof(d));
    spec->ndraws = 2000000;
    spec->index_dist = ADATM_DIST_UNIFORM;
    spec->value_dist = ADATM_VALS_UNIFORM01;
    break;
  }
  case 2: {
    uint64_t const d[5] = {100000, 50000, 20000, 5000, 1000};
    spec->nmodes = 5;
    memcpy(spec->dims, d, sizeof(d));
    spec->ndraws = 100000000;
    spec->index_dist = ADATM_DIST_ZIPF;
    for (int m = 0; m < 5; ++m) spec->zipf_s[m] = 1.1;
    spec->value_dist = ADATM_VALS_LOGNORMAL;
    break;
  }
  case 3: {
    spec->nmodes = 8;
    for (int m = 0; m < 8; ++m) spec->dims[m] = 1000;
    spec->ndraws = 50000000;
    spec->index_dist = ADATM_DIST_CLUSTER;
    spec->nclusters = 100000;
    spec->cluster_sigma = 20.0;
    spec->value_dist = ADATM_VALS_UNIFORM_PM1;
    break;
  }
  default:
    return SPLATT_ERROR_BADINPUT;
  }
  if (ndraws_override > 0) spec->ndraws = ndraws_override;
  return SPLATT_SUCCESS;
}

int adatm_synth(adatm_synth_spec const *spec, adatm_coo *out)
{
  memset(out, 0, sizeof(*out));
  uint32_t const N = spec->nmodes;
  if (N < 1 || N > ADATM_MAX_NMODES || spec->ndraws == 0 ||
      spec->ndraws >= (1ull << 32))
    return SPLATT_ERROR_BADINPUT;
  for (uint32_t m = 0; m < N; ++m)
    if (spec->dims[m] == 0 || spec->dims[m] > (1ull << 31))
      return SPLATT_ERROR_BADINPUT;

  uint64_t const nnz = spec->ndraws;
  out->nmodes = N;
  out->nnz = nnz;
  for (uint32_t m = 0; m < N; ++m) {
    out->dims[m] = spec->dims[m];
    out->ind[m] = (uint32_t *)malloc(nnz * sizeof(uint32_t));
  }
  out->vals = (double *)malloc(nnz * sizeof(double));
  int ok = out->vals != NULL;
  for (uint32_t m = 0; m < N; ++m) ok = ok && out->ind[m] != NULL;
  if (!ok) {
    adatm_coo_free(out);
    return SPLATT_ERROR_NOMEMORY;
  }

  double *cdf[ADATM_MAX_NMODES] = {0};
  double *centres = NULL;
  if (spec->index_dist == ADATM_DIST_ZIPF) {
    for (uint32_t m = 0; m < N; ++m) {
      cdf[m] = zipf_table(spec->dims[m], spec->zipf_s[m]);
      if (!cdf[m]) ok = 0;
    }
  } else if (spec->index_dist == ADATM_DIST_CLUSTER) {
    if (spec->nclusters == 0) ok = 0;
    centres = ok ? (double *)malloc(spec->nclusters * N * sizeof(double))
                 : NULL;
    if (!centres) ok = 0;
    if (ok) {
      #pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < (int64_t)spec->nclusters; ++c) {
        ustream s;
        us_init(&s, spec->seed, (uint64_t)c, TAG_CENTRE);
        for (uint32_t m = 0; m < N; ++m)
          centres[c * N + m] = us_next(&s) * (double)spec->dims[m];
      }
    }
  }
  if (!ok) {
    for (uint32_t m = 0; m < N; ++m) free(cdf[m]);
    free(centres);
    adatm_coo_free(out);
    return SPLATT_ERROR_NOMEMORY;
  }

  #pragma omp parallel for schedule(static, 65536)
  for (int64_t j = 0; j < (int64_t)nnz; ++j) {
    ustream s;
    us_init(&s, spec->seed, (uint64_t)j, TAG_COORD);
    if (spec->index_dist == ADATM_DIST_UNIFORM) {
      for (uint32_t m = 0; m < N; ++m) {
        uint64_t k = (uint64_t)(us_next(&s) * (double)spec->dims[m]);
        if (k >= spec->dims[m]) k = spec->dims[m] - 1;
        out->ind[m][j] = (uint32_t)k;
      }
    } else if (spec->index_dist == ADATM_DIST_ZIPF) {
      for (uint32_t m = 0; m < N; ++m)
        out->ind[m][j] = zipf_draw(cdf[m], spec->dims[m], us_next(&s));
    } else {
      uint64_t c = (uint64_t)(us_next(&s) * (double)spec->nclusters);
      if (c >= spec->nclusters) c = spec->nclusters - 1;
      for (uint32_t m = 0; m < N; ++m) {
        double x = round(centres[c * N + m] +
                         spec->cluster_sigma * us_normal(&s));
        double const hi = (double)(spec->dims[m] - 1);
        if (x < 0.0) x = 0.0;
        if (x > hi) x = hi;
        out->ind[m][j] = (uint32_t)x;
      }
    }
    ustream v;
    us_init(&v, spec->seed, (uint64_t)j, TAG_VALUE);
    double val;
    switch (spec->value_dist) {
    case ADATM_VALS_LOGNORMAL: val = exp(us_normal(&v)); break;
    case ADATM_VALS_UNIFORM_PM1: val = 2.0 * us_next(&v) - 1.0; break;
    default: val = us_next(&v); break;
    }
    out->vals[j] = val;
  }

  for (uint32_t m = 0; m < N; ++m) free(cdf[m]);
  free(centres);
  return SPLATT_SUCCESS;
}

void adatm_rand_matrix(uint64_t seed, uint32_t mode, uint64_t rows,
                       uint64_t rank, double *out)
{
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)rows; ++i) {
    ustream s;
    us_init(&s, seed, (uint64_t)i, TAG_FACTOR + mode);
    for (uint64_t r = 0; r < rank; ++r) out[i * rank + r] = us_next(&s);
  }
}

void adatm_coo_free(adatm_coo *t)
{
  if (!t) return;
  for (uint32_t m = 0; m < ADATM_MAX_NMODES; ++m) {
    free(t->ind[m]);
    t->ind[m] = NULL;
  }
  free(t->vals);
  t->vals = NULL;
  t->nnz = 0;
}
