/*
 * Memoization planner: candidate plans, the analytic cost model and the
 * mode -> (group, tag) map.
 *
 * Restates src/adatm_bench.c:62-508 (decide_product_order,
 * decide_candidate_configs, predict_candidates_adaptive,
 * limit_candidates_by_space, choose_candidates_by_strategy,
 * decide_parameters_auto), src/adatm_rcsf.c:329-406 (predictors),
 * src/adatm_sort.c:10-75 (pair_sort), src/adatm_utils.c:33-51 and 80-98
 * (argmin_elem_range / array_min_range) and src/mttkrp.c:1940-1974 (decide_use_csfs, the
 * TIME_EFFICIENT branch).  The steps are kept identical because the chosen
 * plan, including tie orders of the quicksort, is part of the parity
 * contract.  Deliberate differences:
 *   - N = 3: the reference adds an uninitialised minimum to the candidate
 *     count (adatm_bench.c:303-326); here the count is the number of
 *     candidates the second loop emits (one).
 *   - the chosen position in the list of candidates under the space limit
 *     is mapped back to a candidate index (the reference indexes configs[]
 *     with the position, adatm_bench.c:482-488; identical when no candidate
 *     exceeds the limit).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "adatm_host.h"

int adatm_count_fibers_host(adatm_coo const *t, uint32_t const *order,
                            uint64_t *nfibs);
int adatm_count_fibers_device(adatm_coo const *t, uint32_t const *order,
                              uint64_t *nfibs, int device);

/* ---- quicksort of (mode, dim) pairs; tie order as adatm_sort.c:10-67 ---- */
static void pair_qsort(pair *a, size_t lo, size_t hi)
{
  size_t i = lo + 1, j = hi - 1;
  size_t const mid = lo + (hi - lo) / 2;
  pair const piv = a[mid];
  a[mid] = a[lo];
  while (i < j) {
    if (a[i].y > piv.y) {
      if (a[j].y < piv.y) {
        pair const t = a[i];
        a[i] = a[j];
        a[j] = t;
        ++i;
      }
      --j;
    } else {
      if (a[j].y > piv.y) --j;
      ++i;
    }
  }
  if (a[i].y > piv.y) --i;
  a[lo] = a[i];
  a[i] = piv;
  if (i > lo + 1) pair_qsort(a, lo, i);
  ++i;
  if (hi - i > 1) pair_qsort(a, i, hi);
}

void pair_sort(pair *const pair_list, splatt_idx_t const size)
{
  if (size > 1) pair_qsort(pair_list, 0, (size_t)size);
}

/* first index holding the smallest value > con (adatm_utils.c:33-51) */
splatt_idx_t argmin_elem_range(splatt_idx_t const *const arr,
                               splatt_idx_t const N, splatt_idx_t const con)
{
  splatt_idx_t best = N;
  for (splatt_idx_t i = 0; i < N; ++i)
    if (arr[i] > con && (best == N || arr[i] < arr[best])) best = i;
  return best;
}

/* smallest value > con, 0 when there is none (adatm_utils.c:80-98; the
 * reference leaves the result uninitialised then, D3) */
splatt_idx_t array_min_range(splatt_idx_t const *const arr,
                             splatt_idx_t const N, splatt_idx_t const con)
{
  splatt_idx_t best = 0;
  int found = 0;
  for (splatt_idx_t i = 0; i < N; ++i)
    if (arr[i] > con && (!found || arr[i] < best)) {
      best = arr[i];
      found = 1;
    }
  return best;
}

/* po[m][i] = 1 for factors i > m not yet updated when mode m's MTTKRP runs,
 * po[m][N+i] = 1 for factors i < m already updated (adatm_bench.c:62-74). */
static void product_order_fill(uint32_t N, splatt_idx_t *po /* N x 2N */)
{
  memset(po, 0, (size_t)N * 2 * N * sizeof(splatt_idx_t));
  for (uint32_t m = 0; m < N; ++m) {
    for (uint32_t i = m + 1; i < N; ++i) po[(size_t)m * 2 * N + i] = 1;
    for (uint32_t i = 0; i < m; ++i) po[(size_t)m * 2 * N + N + i] = 1;
  }
}

void decide_product_order(sptensor_t *const tt, splatt_idx_t **product_order)
{
  uint32_t const N = (uint32_t)tt->nmodes;
  for (uint32_t m = 0; m < N; ++m) {
    memset(product_order[m], 0, 2 * N * sizeof(splatt_idx_t));
    for (uint32_t i = m + 1; i < N; ++i) product_order[m][i] = 1;
    for (uint32_t i = 0; i < m; ++i) product_order[m][i + N] = 1;
  }
}

/*
 * Candidate enumeration over a given product order (po row m at
 * po[m*2N ...]).  Writes up to maxc plans; returns the number of plans.
 */
static int enumerate(uint32_t N, uint64_t const *dims, splatt_idx_t const *po,
                     adatm_plan *cands, int maxc)
{
  if (N < 3 || N > ADATM_MAX_NMODES) return 0;
  uint32_t const ngcal = (uint32_t)ceil((double)N / sqrt(2.0 * (N - 1)));
  splatt_idx_t maxim[ADATM_MAX_NMODES + 2];
  splatt_idx_t begin[ADATM_MAX_NMODES + 2];
  splatt_idx_t memo[ADATM_MAX_NMODES + 2];
  splatt_idx_t sum_times[2 * ADATM_MAX_NMODES], tmp_times[2 * ADATM_MAX_NMODES];
  uint32_t tmp_opt[ADATM_MAX_NMODES];
  int cfg = 0;

  for (uint32_t ng = 0; ng < ngcal + 1; ++ng) {
    uint32_t n_grp = ng + 1;
    uint32_t const aver = N / n_grp;
    uint32_t rest = N % n_grp;
    if (aver == 1 && rest > 0) continue;
    if (n_grp >= N) continue;
    memset(maxim, 0, sizeof(maxim));
    memset(begin, 0, sizeof(begin));
    memset(memo, 0, sizeof(memo));
    if (n_grp != 1) {
      for (uint32_t g = 0; g < n_grp; ++g) {
        if (rest > 0) {
          maxim[g] = aver;
          --rest;
        } else {
          maxim[g] = aver - 1;
        }
        begin[g] = maxim[g] + 1;
      }
      memo[0] = 0;
      for (uint32_t g = 1; g < n_grp; ++g)
        memo[g] = memo[g - 1] + maxim[g - 1] + 1;
    } else {
      /* one memoised group plus a plain root group for the last mode */
      maxim[0] = N - 2;
      maxim[1] = 0;
      begin[0] = N - 1;
      begin[1] = 1;
      memo[0] = 0;
      memo[1] = N - 1;
      n_grp = 2;
    }
    splatt_idx_t const nvar = array_min_range(maxim, ngcal + 1, 0);

    for (splatt_idx_t i = 0; i < nvar; ++i) {
      if (cfg >= maxc) return cfg + 1; /* caller's array too small */
      adatm_plan *P = cands + cfg;
      memset(P, 0, sizeof(*P));
      P->n_grp = n_grp;
      for (uint32_t g = 0; g < n_grp; ++g) {
        adatm_group *G = P->grp + g;
        G->memo_mode = (uint32_t)memo[g];
        G->begin_imten = (uint32_t)begin[g];
        G->n_imten = (G->memo_mode != N - 1) ? (uint32_t)(maxim[g] - i) : 0;
        uint32_t const mm = G->memo_mode;
        memset(sum_times, 0, sizeof(sum_times));
        memset(tmp_times, 0, sizeof(tmp_times));
        tmp_opt[0] = mm;
        G->o_opt[0] = mm;
        for (splatt_idx_t j = mm; j <= mm + maxim[g]; ++j)
          for (uint32_t k = 0; k < 2 * N; ++k)
            sum_times[k] += po[(size_t)j * 2 * N + k];
        for (uint32_t k = 0; k < 2 * N; ++k)
          if (po[(size_t)mm * 2 * N + k] > 0) tmp_times[k] = sum_times[k];
        for (uint32_t j = 1; j < N; ++j) {
          splatt_idx_t k = argmin_elem_range(tmp_times, 2 * N, 0);
          tmp_times[k] = 0;
          if (k >= N) k -= N;
          tmp_opt[j] = (uint32_t)k;
        }
        /* the un-memoised tail sorted by dimension (non-decreasing) */
        uint32_t const b = G->begin_imten;
        uint32_t const ns = N - b;
        pair tail[ADATM_MAX_NMODES];
        for (uint32_t j = 0; j < ns; ++j) {
          tail[j].x = tmp_opt[j + b];
          tail[j].y = dims[tmp_opt[j + b]];
        }
        pair_sort(tail, ns);
        for (uint32_t j = 0; j < ns; ++j) G->o_opt[j + b] = (uint32_t)tail[j].x;
        for (uint32_t j = 1; j < b; ++j) G->o_opt[j] = tmp_opt[j];
      }
      ++cfg;
    }
  }
  return cfg;
}

int adatm_plan_enumerate(uint32_t nmodes, uint64_t const *dims,
                         adatm_plan *cands, int maxc)
{
  splatt_idx_t po[ADATM_MAX_NMODES * 2 * ADATM_MAX_NMODES];
  if (nmodes < 3 || nmodes > ADATM_MAX_NMODES) return 0;
  product_order_fill(nmodes, po);
  return enumerate(nmodes, dims, po, cands, maxc);
}

int adatm_plan_candidates(uint32_t nmodes, uint64_t const *dims,
                          adatm_plan *cands, int maxc)
{
  return adatm_plan_enumerate(nmodes, dims, cands, maxc);
}

/* ---- byte / product predictors (adatm_rcsf.c:329-406, reference units) -- */
#define REF_VAL 8u
#define REF_IDX 8u

static uint64_t pred_csf_bytes(uint32_t N, uint64_t nnz, uint64_t const *F)
{
  uint64_t b = nnz * REF_VAL + nnz * REF_IDX;
  for (uint32_t m = 0; m + 1 < N; ++m)
    b += (F[m] + 1) * REF_IDX + F[m] * REF_IDX;
  return b;
}

static uint64_t pred_rcsf_bytes(uint64_t R, uint32_t b, uint32_t nim,
                                uint64_t const *F, int quirks)
{
  uint64_t bytes = 0;
  for (uint32_t i = 0; i < nim; ++i) {
    /* reference counts level b-i; the memo actually stored is level b-1-i */
    uint32_t const lvl = quirks ? b - i : b - 1 - i;
    bytes += F[lvl] * R * REF_VAL;
  }
  return bytes;
}

static uint64_t pred_csf_ops(uint32_t N, uint64_t R, uint64_t const *F)
{
  uint64_t ops = 0;
  for (uint32_t m = N; m >= 1; --m) ops += F[m - 1] * R * 2;
  return ops;
}

static uint64_t pred_rcsf_ops(uint64_t R, uint32_t b, uint32_t nim,
                              uint64_t nre, uint64_t const *F)
{
  uint64_t ops = 0;
  if (nim == 0) return 0;
  uint64_t const end = (uint64_t)b - nim + 1;
  for (uint64_t i = 1; i <= nre; ++i) {
    if (i < end)
      ops += F[end - 1] * R * 2;
    else if (i < b)
      ops += F[i] * R * 2;
  }
  return ops;
}

/* the reference's products-to-reuse count of a group */
static uint64_t num_remttkrp(adatm_plan const *P, uint32_t g, uint32_t N)
{
  if (g + 1 < P->n_grp)
    return P->grp[g + 1].memo_mode - P->grp[g].memo_mode - 1;
  return N - P->grp[g].memo_mode - 1;
}

/* fiber-count cache keyed by level order */
typedef struct {
  uint32_t order[ADATM_MAX_NMODES];
  uint64_t F[ADATM_MAX_NMODES];
} fcache_ent;

static int counts_for(adatm_coo const *t, uint32_t const *order, int device,
                      fcache_ent *cache, int *ncache, uint64_t *F)
{
  uint32_t const N = t->nmodes;
  for (int c = 0; c < *ncache; ++c)
    if (memcmp(cache[c].order, order, N * sizeof(uint32_t)) == 0) {
      memcpy(F, cache[c].F, N * sizeof(uint64_t));
      return SPLATT_SUCCESS;
    }
  int rc = (device >= 0) ? adatm_count_fibers_device(t, order, F, device)
                         : adatm_count_fibers_host(t, order, F);
  if (rc != SPLATT_SUCCESS) return rc;
  if (*ncache < 64) {
    memcpy(cache[*ncache].order, order, N * sizeof(uint32_t));
    memcpy(cache[*ncache].F, F, N * sizeof(uint64_t));
    ++*ncache;
  }
  return SPLATT_SUCCESS;
}

int adatm_count_fibers(adatm_coo const *t, uint32_t const *order,
                       uint64_t *nfibs, int device)
{
  return (device >= 0) ? adatm_count_fibers_device(t, order, nfibs, device)
                       : adatm_count_fibers_host(t, order, nfibs);
}

/*
 * predict_candidates_adaptive (adatm_bench.c:145-272).  With quirks the
 * statistics are refreshed only when the group count changes and the
 * predictors of every group then read the fiber counts of the last group
 * refreshed (nfibs_per_grp), exactly as the reference does.
 */
static int predict(adatm_coo const *t, uint64_t R, adatm_plan *cands, int nc,
                   int quirks, int device)
{
  uint32_t const N = t->nmodes;
  uint64_t const nnz = t->nnz;
  fcache_ent *cache = (fcache_ent *)malloc(64 * sizeof(fcache_ent));
  int ncache = 0;
  if (!cache) return SPLATT_ERROR_NOMEMORY;
  if (device >= 0) {
    /* every order the model will read, counted with one device upload */
    uint32_t *orders = (uint32_t *)malloc(64 * N * sizeof(uint32_t));
    uint32_t no = 0, old = 0;
    for (int c = 0; c < nc && orders; ++c) {
      int const refresh = (c == 0) || !quirks || (cands[c].n_grp != old);
      old = cands[c].n_grp;
      if (!refresh) continue;
      for (uint32_t g = 0; g < cands[c].n_grp && no < 64; ++g) {
        int dup = 0;
        for (uint32_t k = 0; k < no && !dup; ++k)
          dup = !memcmp(orders + k * N, cands[c].grp[g].o_opt, N * sizeof(uint32_t));
        if (!dup) memcpy(orders + (no++) * N, cands[c].grp[g].o_opt, N * sizeof(uint32_t));
      }
    }
    uint64_t *F = (uint64_t *)malloc((size_t)(no ? no : 1) * N * sizeof(uint64_t));
    int rc = (orders && F) ? adatm_count_fibers_device_many(t, orders, no, F, device)
                           : SPLATT_ERROR_NOMEMORY;
    if (rc == SPLATT_SUCCESS)
      for (uint32_t k = 0; k < no; ++k) {
        memcpy(cache[k].order, orders + k * N, N * sizeof(uint32_t));
        memcpy(cache[k].F, F + (size_t)k * N, N * sizeof(uint64_t));
      }
    if (rc == SPLATT_SUCCESS) ncache = (int)no;
    free(orders);
    free(F);
    if (rc != SPLATT_SUCCESS) {
      free(cache);
      return rc;
    }
  }
  uint64_t last_F[ADATM_MAX_NMODES] = {0};
  uint64_t csf_bytes = 0, csf_ops = 0;
  uint32_t old_ngrp = 0;
  int rc = SPLATT_SUCCESS;
  for (int c = 0; c < nc; ++c) {
    adatm_plan *P = cands + c;
    uint64_t rcsf_bytes = 0, rcsf_ops = 0;
    int const refresh = (c == 0) || !quirks || (P->n_grp != old_ngrp);
    if (refresh) {
      old_ngrp = P->n_grp;
      csf_bytes = 0;
      csf_ops = 0;
    }
    for (uint32_t g = 0; g < P->n_grp; ++g) {
      adatm_group const *G = P->grp + g;
      uint64_t F[ADATM_MAX_NMODES];
      if (refresh) {
        rc = counts_for(t, G->o_opt, device, cache, &ncache, F);
        if (rc != SPLATT_SUCCESS) goto out;
        memcpy(last_F, F, sizeof(F));
        csf_bytes += pred_csf_bytes(N, nnz, F);
        csf_ops += pred_csf_ops(N, R, F);
      } else {
        memcpy(F, last_F, sizeof(F));
      }
      uint64_t const nre = num_remttkrp(P, g, N);
      rcsf_bytes += pred_rcsf_bytes(R, G->begin_imten, G->n_imten, F, quirks);
      rcsf_ops += pred_rcsf_ops(R, G->begin_imten, G->n_imten, nre, F);
    }
    P->pspace = csf_bytes + rcsf_bytes;
    P->pops = csf_ops + rcsf_ops;
  }
out:
  free(cache);
  return rc;
}

/* The analytic model of one plan on fresh statistics of every group
 * (Fg[g * N + l]); the reference's model without its statistics reuse. */
void adatm_plan_model(uint32_t N, uint64_t nnz, uint64_t R, adatm_plan const *P,
                      uint64_t const *Fg, uint64_t *pspace, uint64_t *pops)
{
  uint64_t sb = 0, so = 0;
  for (uint32_t g = 0; g < P->n_grp; ++g) {
    adatm_group const *G = P->grp + g;
    uint64_t const *F = Fg + (size_t)g * N;
    sb += pred_csf_bytes(N, nnz, F) +
          pred_rcsf_bytes(R, G->begin_imten, G->n_imten, F, 0);
    so += pred_csf_ops(N, R, F) +
          pred_rcsf_ops(R, G->begin_imten, G->n_imten, num_remttkrp(P, g, N), F);
  }
  *pspace = sb;
  *pops = so;
}

int adatm_plan_analytic(adatm_coo const *t, uint64_t rank, int ref_quirks,
                        double space_limit, int device, adatm_plan *cands,
                        int maxc, int *ncands, int *chosen)
{
  int const nc = adatm_plan_enumerate(t->nmodes, t->dims, cands, maxc);
  if (nc <= 0 || nc > maxc) {
    adatm_set_error("planner: %d candidates for N=%u (capacity %d)", nc,
                    t->nmodes, maxc);
    return SPLATT_ERROR_BADINPUT;
  }
  int rc = predict(t, rank, cands, nc, ref_quirks, device);
  if (rc != SPLATT_SUCCESS) return rc;
  /* limit_candidates_by_space + choose_candidates_by_strategy(1) */
  int best = -1;
  for (int c = 0; c < nc; ++c) {
    if (!((double)cands[c].pspace < space_limit)) continue;
    if (best < 0 || cands[c].pops < cands[best].pops) best = c;
  }
  if (best < 0) {
    adatm_set_error("planner: no candidate under the space limit %.3g B",
                    space_limit);
    return SPLATT_ERROR_NOMEMORY;
  }
  *ncands = nc;
  *chosen = best;
  return SPLATT_SUCCESS;
}

void adatm_plan_use(adatm_plan const *p, uint32_t nmodes, uint32_t *use_grp,
                    uint32_t *use_tag)
{
  for (uint32_t g = 0; g < p->n_grp; ++g) {
    uint32_t const m0 = p->grp[g].memo_mode;
    uint32_t const m1 = (g + 1 < p->n_grp) ? p->grp[g + 1].memo_mode : nmodes;
    use_grp[m0] = g;
    use_tag[m0] = 0;
    for (uint32_t m = m0 + 1; m < m1; ++m) {
      use_grp[m] = g;
      use_tag[m] = 1;
    }
  }
}

/* ---- SPLATT-layout wrappers of the planner ----------------------------- */

static void plan_to_configs(adatm_plan const *src, int nc, uint32_t N,
                            configurations_adaptive *dst)
{
  for (int c = 0; c < nc; ++c) {
    dst[c].n_grp = src[c].n_grp;
    dst[c].pspace = src[c].pspace;
    dst[c].pops = src[c].pops;
    dst[c].ptime = 0;
    dst[c].rtime = 0;
    dst[c].grp_prop =
        (group_properties *)malloc(src[c].n_grp * sizeof(group_properties));
    for (uint32_t g = 0; g < src[c].n_grp; ++g) {
      group_properties *G = dst[c].grp_prop + g;
      G->memo_mode = src[c].grp[g].memo_mode;
      G->begin_imten = src[c].grp[g].begin_imten;
      G->n_imten = src[c].grp[g].n_imten;
      G->o_opt = (splatt_idx_t *)malloc(N * sizeof(splatt_idx_t));
      for (uint32_t m = 0; m < N; ++m) G->o_opt[m] = src[c].grp[g].o_opt[m];
    }
  }
}

splatt_idx_t decide_candidate_configs(sptensor_t *const tt,
                                      splatt_idx_t const **product_order,
                                      configurations_adaptive **configs)
{
  uint32_t const N = (uint32_t)tt->nmodes;
  if (N < 3 || N > ADATM_MAX_NMODES) {
    *configs = NULL;
    return 0;
  }
  splatt_idx_t po[ADATM_MAX_NMODES * 2 * ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < N; ++m)
    for (uint32_t k = 0; k < 2 * N; ++k)
      po[(size_t)m * 2 * N + k] = product_order[m][k];
  uint64_t dims[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < N; ++m) dims[m] = tt->dims[m];
  adatm_plan tmp[256];
  int const nc = enumerate(N, dims, po, tmp, 256);
  *configs = (configurations_adaptive *)malloc(
      (size_t)(nc > 0 ? nc : 1) * sizeof(configurations_adaptive));
  plan_to_configs(tmp, nc, N, *configs);
  return (splatt_idx_t)nc;
}

/* u32 view of a SPLATT coordinate tensor (copies the indices) */
int adatm_coo_from_sptensor(sptensor_t const *tt, adatm_coo *t)
{
  memset(t, 0, sizeof(*t));
  uint32_t const N = (uint32_t)tt->nmodes;
  if (N < 1 || N > ADATM_MAX_NMODES || tt->nnz >= (1ull << 32))
    return SPLATT_ERROR_BADINPUT;
  t->nmodes = N;
  t->nnz = tt->nnz;
  for (uint32_t m = 0; m < N; ++m) {
    if (tt->dims[m] > (1ull << 31)) return SPLATT_ERROR_BADINPUT;
    t->dims[m] = tt->dims[m];
    t->ind[m] = (uint32_t *)malloc(tt->nnz * sizeof(uint32_t));
    if (!t->ind[m]) {
      adatm_coo_free(t);
      return SPLATT_ERROR_NOMEMORY;
    }
    for (uint64_t j = 0; j < tt->nnz; ++j) {
      if (tt->ind[m][j] >= tt->dims[m]) {
        adatm_coo_free(t);
        return SPLATT_ERROR_BADINPUT;
      }
      t->ind[m][j] = (uint32_t)tt->ind[m][j];
    }
  }
  t->vals = (double *)malloc(tt->nnz * sizeof(double));
  if (!t->vals) {
    adatm_coo_free(t);
    return SPLATT_ERROR_NOMEMORY;
  }
  memcpy(t->vals, tt->vals, tt->nnz * sizeof(double));
  return SPLATT_SUCCESS;
}

void predict_candidates_adaptive(sptensor_t *const tt,
                                 splatt_idx_t const nfactors,
                                 configurations_adaptive *configs,
                                 splatt_idx_t const nconfigs)
{
  uint32_t const N = (uint32_t)tt->nmodes;
  adatm_coo t;
  if (adatm_coo_from_sptensor(tt, &t) != SPLATT_SUCCESS) return;
  adatm_plan *P = (adatm_plan *)calloc(nconfigs ? nconfigs : 1,
                                       sizeof(adatm_plan));
  for (splatt_idx_t c = 0; c < nconfigs; ++c) {
    P[c].n_grp = (uint32_t)configs[c].n_grp;
    for (uint32_t g = 0; g < P[c].n_grp; ++g) {
      P[c].grp[g].memo_mode = (uint32_t)configs[c].grp_prop[g].memo_mode;
      P[c].grp[g].begin_imten = (uint32_t)configs[c].grp_prop[g].begin_imten;
      P[c].grp[g].n_imten = (uint32_t)configs[c].grp_prop[g].n_imten;
      for (uint32_t m = 0; m < N; ++m)
        P[c].grp[g].o_opt[m] = (uint32_t)configs[c].grp_prop[g].o_opt[m];
    }
  }
  if (predict(&t, nfactors, P, (int)nconfigs, 1, -1) == SPLATT_SUCCESS)
    for (splatt_idx_t c = 0; c < nconfigs; ++c) {
      configs[c].pspace = P[c].pspace;
      configs[c].pops = P[c].pops;
    }
  free(P);
  adatm_coo_free(&t);
}

splatt_idx_t limit_candidates_by_space(
    configurations_adaptive const *const configs, splatt_idx_t const nconfigs,
    splatt_idx_t **const real_config_locs)
{
  splatt_idx_t n = 0;
  *real_config_locs =
      (splatt_idx_t *)malloc((nconfigs ? nconfigs : 1) * sizeof(splatt_idx_t));
  for (splatt_idx_t c = 0; c < nconfigs; ++c)
    if (configs[c].pspace < SPLATT_ADATM_SLIMIT) (*real_config_locs)[n++] = c;
  return n;
}

/* returns a position in real_config_locs, like the reference */
splatt_idx_t choose_candidates_by_strategy(
    configurations_adaptive const *const configs, splatt_idx_t const nconfigs,
    splatt_idx_t const *const real_config_locs,
    splatt_idx_t const n_real_configs, int strategy)
{
  (void)nconfigs;
  splatt_idx_t best = 0;
  if (n_real_configs == 0) return 0;
  size_t best_ops = configs[real_config_locs[0]].pops;
  size_t best_space = configs[real_config_locs[0]].pspace;
  for (splatt_idx_t c = 1; c < n_real_configs; ++c) {
    configurations_adaptive const *C = configs + real_config_locs[c];
    if (strategy == 1 && C->pops < best_ops) {
      best_ops = C->pops;
      best = c;
    } else if (strategy == 2 && C->pspace < best_space) {
      best_space = C->pspace;
      best = c;
    }
  }
  return best;
}

splatt_idx_t decide_parameters_auto(sptensor_t *const tt,
                                    splatt_idx_t const nfactors, int strategy,
                                    splatt_idx_t const **product_order,
                                    group_properties **grp_prop)
{
  uint32_t const N = (uint32_t)tt->nmodes;
  configurations_adaptive *configs = NULL;
  splatt_idx_t const nc = decide_candidate_configs(tt, product_order, &configs);
  printf("Number of configurations: %lu\n", (unsigned long)nc);
  fflush(stdout);
  predict_candidates_adaptive(tt, nfactors, configs, nc);
  splatt_idx_t *locs = NULL;
  splatt_idx_t const nreal = limit_candidates_by_space(configs, nc, &locs);
  splatt_idx_t pos = choose_candidates_by_strategy(configs, nc, locs, nreal,
                                                   strategy);
  splatt_idx_t const idx = nreal ? locs[pos] : 0;
  splatt_idx_t const n_grp = configs[idx].n_grp;
  *grp_prop = (group_properties *)malloc(n_grp * sizeof(group_properties));
  for (splatt_idx_t g = 0; g < n_grp; ++g) {
    (*grp_prop)[g] = configs[idx].grp_prop[g];
    (*grp_prop)[g].o_opt = (splatt_idx_t *)malloc(N * sizeof(splatt_idx_t));
    memcpy((*grp_prop)[g].o_opt, configs[idx].grp_prop[g].o_opt,
           N * sizeof(splatt_idx_t));
  }
  for (splatt_idx_t c = 0; c < nc; ++c) {
    for (splatt_idx_t g = 0; g < configs[c].n_grp; ++g)
      free(configs[c].grp_prop[g].o_opt);
    free(configs[c].grp_prop);
  }
  free(configs);
  free(locs);
  return n_grp;
}

/*
 * SPACE_EFFICIENT allocation (adatm_rcsf.c:528-552, mttkrp.c:1976-2029): only
 * the groups that memoise keep a CSF; a mode whose group stores nothing is
 * computed from the leaves ("tag 2") of the CSF of the nearest memoising group
 * before it.  The reference locates that group with locate_ind_special, which
 * stops after comparing two entries (D7, adatm_utils.c:17-24), and skips the
 * last group when there is one (D8); here every group is located.
 */
int adatm_plan_space_efficient(adatm_plan const *p, uint32_t N, adatm_plan *out,
                               uint32_t *use_grp, uint32_t *use_tag)
{
  int keep[ADATM_MAX_NMODES];
  uint32_t nk = 0;
  for (uint32_t g = 0; g < p->n_grp; ++g) keep[g] = p->grp[g].n_imten > 0;
  for (uint32_t g = 0; g < p->n_grp; ++g) nk += keep[g] ? 1u : 0u;
  if (nk == 0 || p->n_grp == 1) { /* one CSF either way (adatm_rcsf.c:513-516) */
    for (uint32_t g = 0; g < p->n_grp; ++g) keep[g] = (g == 0);
    nk = 1;
  }
  memset(out, 0, sizeof(*out));
  int32_t idx[ADATM_MAX_NMODES];
  for (uint32_t g = 0; g < p->n_grp; ++g) {
    idx[g] = -1;
    if (keep[g]) {
      idx[g] = (int32_t)out->n_grp;
      out->grp[out->n_grp++] = p->grp[g];
    }
  }
  out->pspace = p->pspace;
  out->pops = p->pops;
  for (uint32_t g = 0; g < p->n_grp; ++g) {
    uint32_t const m0 = p->grp[g].memo_mode;
    uint32_t const m1 = (g + 1 < p->n_grp) ? p->grp[g + 1].memo_mode : N;
    int32_t c = idx[g];
    uint32_t tag0 = 0, tagr = 1;
    if (c < 0) { /* the nearest memoising group before g (else the first) */
      for (int32_t h = (int32_t)g - 1; h >= 0 && c < 0; --h) c = idx[h];
      if (c < 0) c = 0;
      tag0 = tagr = 2;
    }
    use_grp[m0] = (uint32_t)c;
    use_tag[m0] = tag0;
    for (uint32_t m = m0 + 1; m < m1; ++m) {
      use_grp[m] = (uint32_t)c;
      use_tag[m] = tagr;
    }
  }
  return (int)nk;
}

void decide_use_csfs(splatt_idx_t const nmodes,
                     group_properties *const grp_prop, int const n_grp,
                     splatt_idx_t const n_csf, splatt_idx_t *use_csfs,
                     splatt_idx_t *use_tags)
{
  if ((int)n_csf != n_grp) { /* SPACE_EFFICIENT: CSFs of memoising groups only */
    adatm_plan p, q;
    uint32_t ug[ADATM_MAX_NMODES], ut[ADATM_MAX_NMODES];
    memset(&p, 0, sizeof(p));
    p.n_grp = (uint32_t)n_grp;
    for (int g = 0; g < n_grp && g < ADATM_MAX_NMODES; ++g) {
      p.grp[g].memo_mode = (uint32_t)grp_prop[g].memo_mode;
      p.grp[g].begin_imten = (uint32_t)grp_prop[g].begin_imten;
      p.grp[g].n_imten = (uint32_t)grp_prop[g].n_imten;
    }
    adatm_plan_space_efficient(&p, (uint32_t)nmodes, &q, ug, ut);
    for (splatt_idx_t m = 0; m < nmodes; ++m) {
      use_csfs[m] = ug[m];
      use_tags[m] = ut[m];
    }
    return;
  }
  for (int g = 0; g < n_grp; ++g) {
    splatt_idx_t const m0 = grp_prop[g].memo_mode;
    splatt_idx_t const m1 = (g + 1 < n_grp) ? grp_prop[g + 1].memo_mode : nmodes;
    use_csfs[m0] = (splatt_idx_t)g;
    use_tags[m0] = 0;
    for (splatt_idx_t m = m0 + 1; m < m1; ++m) {
      use_csfs[m] = (splatt_idx_t)g;
      use_tags[m] = 1;
    }
  }
}

/* ---- storage / reverse order helpers ----------------------------------- */

void rcsf_reverse_mode_order(splatt_idx_t const begin, splatt_idx_t const mode,
                             splatt_idx_t *const rdims)
{
  splatt_idx_t j = 0;
  for (int64_t i = (int64_t)begin; i >= 0; --i)
    if ((splatt_idx_t)i != mode) rdims[j++] = (splatt_idx_t)i;
}

size_t predict_csf_bytes_adaptive(sptensor_t *const tt,
                                  splatt_idx_t const *const nfibs_per_grp)
{
  uint64_t F[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < tt->nmodes && m < ADATM_MAX_NMODES; ++m)
    F[m] = nfibs_per_grp[m];
  return pred_csf_bytes((uint32_t)tt->nmodes, tt->nnz, F);
}

size_t predict_rcsf_bytes_adaptive(sptensor_t *const tt,
                                   splatt_idx_t const nfactors,
                                   splatt_idx_t const begin_imten,
                                   splatt_idx_t const n_imten,
                                   splatt_idx_t const *const nfibs_per_grp)
{
  uint64_t F[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < tt->nmodes && m < ADATM_MAX_NMODES; ++m)
    F[m] = nfibs_per_grp[m];
  return pred_rcsf_bytes(nfactors, (uint32_t)begin_imten, (uint32_t)n_imten,
                         F, 1);
}

size_t predict_csf_ops_adaptive(sptensor_t *const tt,
                                splatt_idx_t const nfactors,
                                splatt_idx_t const *const nfibs_per_grp)
{
  uint64_t F[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < tt->nmodes && m < ADATM_MAX_NMODES; ++m)
    F[m] = nfibs_per_grp[m];
  return pred_csf_ops((uint32_t)tt->nmodes, nfactors, F);
}

size_t predict_rcsf_ops_adaptive(sptensor_t *const tt,
                                 splatt_idx_t const nfactors,
                                 splatt_idx_t const begin_imten,
                                 splatt_idx_t const n_imten,
                                 splatt_idx_t const num_reMTTKRP,
                                 splatt_idx_t const *const nfibs_per_grp)
{
  uint64_t F[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < tt->nmodes && m < ADATM_MAX_NMODES; ++m)
    F[m] = nfibs_per_grp[m];
  return pred_rcsf_ops(nfactors, (uint32_t)begin_imten, (uint32_t)n_imten,
                       num_reMTTKRP, F);
}

/*
 * count_nfibs (adatm_rcsf.c:410-500): tt sorted by o_opt; level `mode`'s
 * fiber count from the previous level's boundaries `fprev` (rewritten to
 * child indices, as the reference does) into a new boundary array *fnext.
 */
void count_nfibs(sptensor_t const *const tt, splatt_idx_t const mode,
                 splatt_idx_t const *const o_opt, splatt_idx_t *const fprev,
                 splatt_idx_t *const nfibs_per_grp, splatt_idx_t **const fnext)
{
  splatt_idx_t const nnz = tt->nnz;
  splatt_idx_t const *ix = tt->ind[o_opt[mode]];
  if (mode == 0) {
    splatt_idx_t nf = 1;
    for (splatt_idx_t x = 1; x < nnz; ++x) nf += (ix[x] != ix[x - 1]);
    nfibs_per_grp[0] = nf;
    splatt_idx_t *fn = (splatt_idx_t *)malloc((nf + 1) * sizeof(splatt_idx_t));
    splatt_idx_t k = 0;
    fn[k++] = 0;
    for (splatt_idx_t x = 1; x < nnz; ++x)
      if (ix[x] != ix[x - 1]) fn[k++] = x;
    fn[nf] = nnz;
    *fnext = fn;
    return;
  }
  splatt_idx_t const np = nfibs_per_grp[mode - 1];
  splatt_idx_t nf = 0;
  for (splatt_idx_t s = 0; s < np; ++s) {
    ++nf;
    for (splatt_idx_t f = fprev[s] + 1; f < fprev[s + 1]; ++f)
      nf += (ix[f] != ix[f - 1]);
  }
  nfibs_per_grp[mode] = nf;
  splatt_idx_t *fn = (splatt_idx_t *)malloc((nf + 1) * sizeof(splatt_idx_t));
  splatt_idx_t k = 0;
  for (splatt_idx_t s = 0; s < np; ++s) {
    splatt_idx_t const st = fprev[s] + 1, en = fprev[s + 1];
    fprev[s] = k;
    fn[k++] = st - 1;
    for (splatt_idx_t f = st; f < en; ++f)
      if (ix[f] != ix[f - 1]) fn[k++] = f;
  }
  fprev[np] = nf;
  fn[nf] = nnz;
  *fnext = fn;
}
