/*
 * Index relabelling of a coordinate tensor: SPLATT's tt_perm / perm_apply /
 * perm_matrix family (reorder.h and reorder.c of SPLATT v1.1.1, which the
 * reference patches but does not ship).  The reference reaches it from
 * `splatt bench -z TYPE -p FILE [-m MODE]` (cmd_bench.c:78-85, 119-149,
 * 243-251): the tensor is relabelled once before the timed runs, the factor
 * matrices follow their rows to the new labels (p_shuffle_mats,
 * adatm_bench.c:40-51, 529) and are moved back afterwards
 * (adatm_bench.c:654).  Every MTTKRP in original labels is therefore the
 * same as without relabelling; only the CSF layout (and so the order of
 * the factor-row gathers) changes.
 *
 * The partition file holds one part id per vertex.  The vertices:
 *   PERM_GRAPH     the indices of every mode (mode 0's first): the
 *                  m-partite graph SPLATT partitions with METIS;
 *   PERM_HGRAPH    the fibers (level N-2 nodes) of the CSF rooted at the
 *                  basis mode: the fiber hypergraph SPLATT partitions with
 *                  PaToH;
 *   PERM_FIBSCHED  the slices (level 0 nodes) of the same CSF: a schedule
 *                  of root slices over parts.
 * PERM_GRAPH orders each mode's indices by (part, index).  The other two
 * visit their vertices by (part, CSF position) and hand out each mode's
 * new labels in the order the visit first touches the indices.  The exact
 * tie-breaking of SPLATT's reorder.c cannot be checked here (the file is
 * absent); results in original labels do not depend on it.
 *
 * The CSF of the basis mode is built by the GPU radix sort and scans when
 * a device is given (the host build otherwise); the labelling itself is a
 * single ordered pass over the CSF's levels.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../../include/adatm.h"
#include "../../../include/splatt.h"
#include "adatm_host.h"

#define NONE 0xffffffffu

int adatm_part_read(char const *fname, uint64_t *count, uint32_t **parts)
{
  *count = 0;
  *parts = NULL;
  FILE *f = fopen(fname, "r");
  if (!f) {
    adatm_set_error("cannot open partition file '%s'", fname);
    return SPLATT_ERROR_BADINPUT;
  }
  uint64_t cap = 1024, n = 0;
  uint32_t *p = (uint32_t *)malloc(cap * sizeof(uint32_t));
  char tok[64];
  int rc = SPLATT_SUCCESS;
  while (p && fscanf(f, "%63s", tok) == 1) {
    char *end = NULL;
    errno = 0;
    unsigned long long const v = strtoull(tok, &end, 10);
    if (tok[0] == '-' || *end != '\0' || errno || v >= NONE) {
      adatm_set_error("partition file '%s': entry %llu ('%s') is not a part id", fname,
                      (unsigned long long)n + 1, tok);
      rc = SPLATT_ERROR_BADINPUT;
      break;
    }
    if (n == cap) {
      cap *= 2;
      uint32_t *q = (uint32_t *)realloc(p, cap * sizeof(uint32_t));
      if (!q) {
        free(p);
        p = NULL;
        break;
      }
      p = q;
    }
    p[n++] = (uint32_t)v;
  }
  fclose(f);
  if (!p) {
    adatm_set_error("partition file '%s': out of memory", fname);
    return SPLATT_ERROR_NOMEMORY;
  }
  if (rc != SPLATT_SUCCESS) {
    free(p);
    return rc;
  }
  *count = n;
  *parts = p;
  return SPLATT_SUCCESS;
}

/* the basis CSF's level order: the basis mode, then the others shortest
 * first (ties by mode), SPLATT's order for a mode's own CSF */
static void basis_order(adatm_coo const *t, uint32_t mode, uint32_t *order)
{
  uint32_t k = 0, N = t->nmodes;
  order[k++] = mode;
  for (uint32_t m = 0; m < N; ++m)
    if (m != mode) order[k++] = m;
  for (uint32_t i = 2; i < N; ++i)
    for (uint32_t j = i; j > 1 && t->dims[order[j]] < t->dims[order[j - 1]]; --j) {
      uint32_t const x = order[j];
      order[j] = order[j - 1];
      order[j - 1] = x;
    }
}

static int check_args(adatm_coo const *t, int type, uint32_t mode)
{
  if (!t || t->nmodes < 2 || t->nmodes > ADATM_MAX_NMODES || t->nnz == 0) {
    adatm_set_error("relabel: bad tensor");
    return SPLATT_ERROR_BADINPUT;
  }
  if (type != PERM_GRAPH && type != PERM_HGRAPH && type != PERM_FIBSCHED) {
    adatm_set_error("relabel: unknown reordering type %d", type);
    return SPLATT_ERROR_BADINPUT;
  }
  if (type != PERM_GRAPH && mode >= t->nmodes) {
    adatm_set_error("relabel: basis mode %u of a %u-mode tensor", mode + 1, t->nmodes);
    return SPLATT_ERROR_BADINPUT;
  }
  return SPLATT_SUCCESS;
}

int adatm_perm_nvtxs(adatm_coo const *t, int type, uint32_t mode, int device,
                     uint64_t *nvtxs)
{
  int rc = check_args(t, type, mode);
  if (rc != SPLATT_SUCCESS) return rc;
  if (type == PERM_GRAPH) {
    uint64_t s = 0;
    for (uint32_t m = 0; m < t->nmodes; ++m) s += t->dims[m];
    *nvtxs = s;
    return SPLATT_SUCCESS;
  }
  uint32_t order[ADATM_MAX_NMODES];
  uint64_t F[ADATM_MAX_NMODES];
  basis_order(t, mode, order);
  rc = adatm_count_fibers(t, order, F, device);
  if (rc == SPLATT_SUCCESS) *nvtxs = type == PERM_HGRAPH ? F[t->nmodes - 2] : F[0];
  return rc;
}

/* vertices 0..n-1 ordered by (part, vertex): a stable counting sort */
static uint32_t *order_by_part(uint32_t const *part, uint64_t n)
{
  uint32_t mx = 0;
  for (uint64_t v = 0; v < n; ++v) mx = part[v] > mx ? part[v] : mx;
  uint64_t *start = (uint64_t *)calloc((size_t)mx + 2, sizeof(uint64_t));
  uint32_t *ord = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  if (!start || !ord) {
    free(start);
    free(ord);
    return NULL;
  }
  for (uint64_t v = 0; v < n; ++v) start[part[v] + 1]++;
  for (uint64_t p = 0; p <= mx; ++p) start[p + 1] += start[p];
  for (uint64_t v = 0; v < n; ++v) ord[start[part[v]]++] = (uint32_t)v;
  free(start);
  return ord;
}

static int perm_graph(adatm_coo const *t, uint32_t const *parts, uint32_t **perms)
{
  uint64_t off = 0;
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    uint32_t *ord = order_by_part(parts + off, t->dims[m]);
    if (!ord) return SPLATT_ERROR_NOMEMORY;
    for (uint64_t k = 0; k < t->dims[m]; ++k) perms[m][ord[k]] = (uint32_t)k;
    free(ord);
    off += t->dims[m];
  }
  return SPLATT_SUCCESS;
}

/* first-touch labels over the vertices of level L of the CSF h */
static int perm_first_touch(adatm_coo const *t, adatm_hcsf const *h, uint32_t L,
                            uint32_t const *parts, uint32_t **perms)
{
  uint32_t const N = h->N;
  uint64_t const nv = h->F[L];
  uint32_t *ord = order_by_part(parts, nv);
  uint32_t *par[ADATM_MAX_NMODES] = {0};
  int rc = SPLATT_ERROR_NOMEMORY;
  if (!ord) goto out;
  /* parents of the path levels 1..L from the child pointers */
  for (uint32_t l = 1; l <= L; ++l) {
    par[l] = (uint32_t *)malloc((h->F[l] ? h->F[l] : 1) * sizeof(uint32_t));
    if (!par[l]) goto out;
    uint32_t const *fp = h->fptr[l - 1];
    for (uint64_t p = 0; p < h->F[l - 1]; ++p)
      for (uint32_t c = fp[p]; c < fp[p + 1]; ++c) par[l][c] = (uint32_t)p;
  }
  uint32_t next[ADATM_MAX_NMODES] = {0};
  for (uint32_t m = 0; m < N; ++m)
    for (uint64_t i = 0; i < t->dims[m]; ++i) perms[m][i] = NONE;
  for (uint64_t k = 0; k < nv; ++k) {
    uint32_t const v = ord[k];
    /* the vertex's path, root first */
    uint32_t node[ADATM_MAX_NMODES];
    node[L] = v;
    for (uint32_t l = L; l >= 1; --l) node[l - 1] = par[l][node[l]];
    for (uint32_t l = 0; l <= L; ++l) {
      uint32_t *lab = perms[h->order[l]];
      uint32_t const id = h->fids[l][node[l]];
      if (lab[id] == NONE) lab[id] = next[l]++;
    }
    /* its subtree, level by level in CSF order */
    uint64_t lo = v, hi = (uint64_t)v + 1;
    for (uint32_t l = L + 1; l < N; ++l) {
      lo = h->fptr[l - 1][lo];
      hi = h->fptr[l - 1][hi];
      uint32_t *lab = perms[h->order[l]];
      uint32_t const *ids = h->fids[l];
      for (uint64_t x = lo; x < hi; ++x)
        if (lab[ids[x]] == NONE) lab[ids[x]] = next[l]++;
    }
  }
  /* indices without a nonzero keep their relative order after the others */
  for (uint32_t l = 0; l < N; ++l) {
    uint32_t const m = h->order[l];
    for (uint64_t i = 0; i < t->dims[m]; ++i)
      if (perms[m][i] == NONE) perms[m][i] = next[l]++;
  }
  rc = SPLATT_SUCCESS;
out:
  free(ord);
  for (uint32_t l = 0; l < ADATM_MAX_NMODES; ++l) free(par[l]);
  return rc;
}

int adatm_perm_compute(adatm_coo const *t, int type, uint32_t mode,
                       uint32_t const *parts, uint64_t nparts, int device,
                       uint32_t **perms)
{
  int rc = check_args(t, type, mode);
  if (rc != SPLATT_SUCCESS) return rc;
  for (uint32_t m = 0; m < t->nmodes; ++m)
    if (t->dims[m] >= NONE) {
      adatm_set_error("relabel: mode %u has %llu indices", m + 1,
                      (unsigned long long)t->dims[m]);
      return SPLATT_ERROR_BADINPUT;
    }
  if (type == PERM_GRAPH) {
    uint64_t nv = 0;
    for (uint32_t m = 0; m < t->nmodes; ++m) nv += t->dims[m];
    if (nparts < nv) {
      adatm_set_error("partition file has %llu entries; the graph has %llu vertices",
                      (unsigned long long)nparts, (unsigned long long)nv);
      return SPLATT_ERROR_BADINPUT;
    }
    return perm_graph(t, parts, perms);
  }
  uint32_t order[ADATM_MAX_NMODES];
  basis_order(t, mode, order);
  adatm_hcsf h;
  rc = device >= 0 ? adatm_hcsf_build_device(t, order, device, &h)
                   : adatm_hcsf_build(t, order, &h);
  if (rc != SPLATT_SUCCESS) return rc;
  uint32_t const L = type == PERM_HGRAPH ? t->nmodes - 2 : 0;
  if (nparts < h.F[L]) {
    adatm_set_error("partition file has %llu entries; mode %u's CSF has %llu %s",
                    (unsigned long long)nparts, mode + 1, (unsigned long long)h.F[L],
                    type == PERM_HGRAPH ? "fibers" : "slices");
    rc = SPLATT_ERROR_BADINPUT;
  } else {
    rc = perm_first_touch(t, &h, L, parts, perms);
  }
  adatm_hcsf_free(&h);
  return rc;
}

void adatm_coo_relabel(adatm_coo *t, uint32_t *const *perms)
{
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    uint32_t const *p = perms[m];
    if (!p) continue;
    uint32_t *ix = t->ind[m];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)t->nnz; ++j) ix[j] = p[ix[j]];
  }
}

/* ---- SPLATT's interface ---------------------------------------------------- */
permutation_t *perm_alloc(splatt_idx_t const *const dims, splatt_idx_t const nmodes)
{
  permutation_t *p = (permutation_t *)calloc(1, sizeof(permutation_t));
  if (!p) return NULL;
  for (splatt_idx_t m = 0; m < nmodes && m < SPLATT_MAX_NMODES; ++m) {
    splatt_idx_t const n = dims[m] ? dims[m] : 1;
    p->perms[m] = (splatt_idx_t *)malloc(n * sizeof(splatt_idx_t));
    p->iperms[m] = (splatt_idx_t *)malloc(n * sizeof(splatt_idx_t));
    if (!p->perms[m] || !p->iperms[m]) {
      perm_free(p);
      return NULL;
    }
    for (splatt_idx_t i = 0; i < dims[m]; ++i) p->perms[m][i] = p->iperms[m][i] = i;
  }
  return p;
}

void perm_free(permutation_t *perm)
{
  if (!perm) return;
  for (int m = 0; m < SPLATT_MAX_NMODES; ++m) {
    free(perm->perms[m]);
    free(perm->iperms[m]);
  }
  free(perm);
}

void perm_apply(sptensor_t *const tt, splatt_idx_t **perm)
{
  for (splatt_idx_t m = 0; m < tt->nmodes; ++m) {
    splatt_idx_t const *p = perm[m];
    if (!p) continue;
    splatt_idx_t *ix = tt->ind[m];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)tt->nnz; ++j) ix[j] = p[ix[j]];
  }
}

matrix_t *perm_matrix(matrix_t const *const mat, splatt_idx_t const *const perm,
                      matrix_t *retmat)
{
  splatt_idx_t const I = mat->I, J = mat->J;
  if (!retmat) {
    retmat = mat_alloc(I, J);
    if (!retmat) return NULL;
    retmat->rowmajor = mat->rowmajor;
  }
  if (mat->rowmajor) {
    for (splatt_idx_t i = 0; i < I; ++i)
      memcpy(retmat->vals + perm[i] * J, mat->vals + i * J, J * sizeof(splatt_val_t));
  } else {
    for (splatt_idx_t j = 0; j < J; ++j)
      for (splatt_idx_t i = 0; i < I; ++i)
        retmat->vals[j * I + perm[i]] = mat->vals[j * I + i];
  }
  return retmat;
}

permutation_t *tt_perm(sptensor_t *const tt, splatt_perm_type const type,
                       splatt_idx_t const mode, char const *const pfile)
{
  if (!tt || tt->nmodes < 2 || tt->nmodes > ADATM_MAX_NMODES || !pfile) {
    adatm_set_error("tt_perm: bad arguments");
    return NULL;
  }
  uint32_t const N = (uint32_t)tt->nmodes;
  adatm_coo c;
  uint64_t dims[ADATM_MAX_NMODES];
  for (uint32_t m = 0; m < N; ++m) {
    if (tt->dims[m] >= NONE) {
      adatm_set_error("tt_perm: mode %u has %llu indices", m + 1,
                      (unsigned long long)tt->dims[m]);
      return NULL;
    }
    dims[m] = tt->dims[m];
  }
  if (adatm_coo_alloc(N, tt->nnz, dims, &c) != SPLATT_SUCCESS) return NULL;
  for (uint32_t m = 0; m < N; ++m)
    for (splatt_idx_t j = 0; j < tt->nnz; ++j) c.ind[m][j] = (uint32_t)tt->ind[m][j];
  memcpy(c.vals, tt->vals, tt->nnz * sizeof(double));
  uint64_t np = 0;
  uint32_t *parts = NULL;
  uint32_t *pm[ADATM_MAX_NMODES] = {0};
  permutation_t *out = NULL;
  int rc = adatm_part_read(pfile, &np, &parts);
  for (uint32_t m = 0; rc == SPLATT_SUCCESS && m < N; ++m)
    if (!(pm[m] = (uint32_t *)malloc((dims[m] ? dims[m] : 1) * sizeof(uint32_t))))
      rc = SPLATT_ERROR_NOMEMORY;
  /* the basis CSF on the first GPU */
  if (rc == SPLATT_SUCCESS) rc = adatm_perm_compute(&c, (int)type, (uint32_t)mode, parts,
                                                    np, 0, pm);
  if (rc == SPLATT_SUCCESS) out = perm_alloc(tt->dims, N);
  if (out) {
    for (uint32_t m = 0; m < N; ++m)
      for (uint64_t i = 0; i < dims[m]; ++i) {
        out->perms[m][i] = pm[m][i];
        out->iperms[m][pm[m][i]] = i;
      }
    perm_apply(tt, out->perms);
  }
  for (uint32_t m = 0; m < N; ++m) free(pm[m]);
  free(parts);
  adatm_coo_free(&c);
  return out;
}
