/*
 * Oracle, C half -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's
 * cpu_baseline leg).  Nothing in the product links or calls this file.
 *
 * FP64 C99 + OpenMP restatement of the reference AdaTM CPU path for one
 * memoization plan, race-free:
 *   - CSF per group from coordinates: lexicographic sort by the group's level
 *     order (tt_sort) and one node per distinct prefix
 *     (p_mk_outerptr / p_mk_fptr, src/adatm_rcsf.c:21-209);
 *   - memoised root MTTKRP: leaf SpTTM then one Hadamard-reduce pass per
 *     level, each an omp for schedule(dynamic,16) over parents
 *     (p_spttm / p_spt2t_add_hada / p_spt2m_add_hada, src/mttkrp.c:828-928,
 *     orchestration src/mttkrp.c:934-1079), storing levels b-1..b-n_imten.
 *     Memo buffers are overwritten, not accumulated into (reference defect:
 *     they are zeroed once, src/adatm_rcsf.c:273-274, then += every call);
 *     phases are separated by the implicit barrier of `omp for` (reference:
 *     `nowait` phases race, src/mttkrp.c:840/875/910);
 *   - plain root MTTKRP (groups that store nothing): per-slice depth-first
 *     traversal (p_csf_mttkrp_root + p_propagate_up, src/mttkrp.c:241-304,
 *     766-824);
 *   - reuse MTTKRP of a level d: per-slice depth-first traversal carrying
 *     Y = prod of ancestor factor rows, Z_d from the memo directly or
 *     propagated up from memo level L (p_propagate_up_reuse), output rows
 *     updated under striped locks (p_csf_mttkrp_internal_reuse_adaptive,
 *     src/mttkrp.c:169-236, 1547-1682; the 1024 locks are initialised here);
 *   - CP-ALS: cpd_als_iterate_adaptive (src/adatm_cpd.c:256-374) with Gram
 *     matrices, (hadamard of Grams)^-1 by Cholesky, A = M V^-1, column
 *     normalisation (2-norm on the first sweep, max(1, column max) after)
 *     and the fit of src/adatm_cpd.c:59-200.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define OMAXN 16
#define NLOCKS 1024

typedef struct {
  uint32_t memo, b, nim;
  uint32_t order[OMAXN];
  uint64_t F[OMAXN];
  uint64_t *fptr[OMAXN]; /* l < N-1 */
  uint64_t *fids[OMAXN]; /* all levels */
  double *vals;
  double *memo_vals[OMAXN]; /* level -> F_l x R */
} ogroup;

typedef struct {
  uint32_t N, R, ngrp;
  uint64_t nnz;
  uint64_t dims[OMAXN];
  ogroup g[OMAXN];
  uint32_t use_grp[OMAXN], use_tag[OMAXN], level_of[OMAXN][OMAXN];
  double *tmp[2];
  omp_lock_t locks[NLOCKS];
  double ttnormsq;
} ohandle;

/* ---- CSF construction ---------------------------------------------------- */
/* Stable LSD radix sort of the positions 0..n-1 by the keys
 * keys[0], ..., keys[nk-1] (keys[0] most significant; key k < kdim[k]):
 * 11-bit digits, each pass a per-thread histogram over a contiguous range,
 * bucket offsets in (digit, thread) order and an in-order scatter, so the
 * result is the unique lexicographic order with ties kept in input order
 * (SPLATT's tt_sort leaves coordinates in lexicographic order by the level
 * order, src/adatm_rcsf.c:179). */
#define RDIG 11
static void radix_lex(uint64_t n, uint32_t nk, uint32_t const *const *keys,
                      uint64_t const *kdim, uint32_t *perm)
{
  uint32_t *tmp = (uint32_t *)malloc(n * sizeof(uint32_t));
  int const T = omp_get_max_threads();
  uint64_t *hist = (uint64_t *)malloc((size_t)T * (1u << RDIG) * sizeof(uint64_t));
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j) perm[j] = (uint32_t)j;
  uint32_t *src = perm, *dst = tmp;
  for (int k = (int)nk - 1; k >= 0; --k) {
    uint32_t bits = 0;
    while (bits < 32 && (kdim[k] - 1) >> bits) ++bits;
    uint32_t const *key = keys[k];
    for (uint32_t sh = 0; sh < bits; sh += RDIG) {
      uint32_t const mask = (1u << RDIG) - 1u;
      #pragma omp parallel num_threads(T)
      {
        int const t = omp_get_thread_num(), nt = omp_get_num_threads();
        uint64_t const lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
        uint64_t *h = hist + (size_t)t * (1u << RDIG);
        memset(h, 0, (1u << RDIG) * sizeof(uint64_t));
        for (uint64_t j = lo; j < hi; ++j) h[(key[src[j]] >> sh) & mask]++;
        #pragma omp barrier
        #pragma omp single
        {
          uint64_t run = 0;
          for (uint32_t d = 0; d <= mask; ++d)
            for (int q = 0; q < nt; ++q) {
              uint64_t const c = hist[(size_t)q * (1u << RDIG) + d];
              hist[(size_t)q * (1u << RDIG) + d] = run;
              run += c;
            }
        }
        for (uint64_t j = lo; j < hi; ++j) {
          uint32_t const v = src[j];
          dst[h[(key[v] >> sh) & mask]++] = v;
        }
      }
      uint32_t *x = src;
      src = dst;
      dst = x;
    }
  }
  if (src != perm) memcpy(perm, src, n * sizeof(uint32_t));
  free(hist);
  free(tmp);
}

static void build_group(ohandle *h, ogroup *g, uint32_t const *const *ind,
                        double const *vals)
{
  uint32_t const N = h->N;
  uint64_t const n = h->nnz;
  uint32_t const *key[OMAXN];
  uint64_t kdim[OMAXN];
  for (uint32_t k = 0; k < N; ++k) {
    key[k] = ind[g->order[k]];
    kdim[k] = h->dims[g->order[k]];
  }
  uint32_t *perm = (uint32_t *)malloc(n * sizeof(uint32_t));
  radix_lex(n, N, key, kdim, perm);
  g->vals = (double *)malloc(n * sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j) g->vals[j] = vals[perm[j]];
  /* first differing level of consecutive sorted coordinates */
  uint8_t *fd = (uint8_t *)malloc(n + 1);
  fd[0] = 0;
  #pragma omp parallel for schedule(static)
  for (int64_t j = 1; j < (int64_t)n; ++j) {
    uint8_t l = (uint8_t)N;
    for (uint32_t k = 0; k < N; ++k)
      if (key[k][perm[j]] != key[k][perm[j - 1]]) {
        l = (uint8_t)k;
        break;
      }
    fd[j] = l;
  }
  /* node starts per level */
  uint64_t *start[OMAXN];
  for (uint32_t l = 0; l < N; ++l) {
    uint64_t c = 0;
    for (uint64_t j = 0; j < n; ++j) c += (j == 0 || fd[j] <= l);
    g->F[l] = (l == N - 1) ? n : c;
    start[l] = (uint64_t *)malloc((g->F[l] + 1) * sizeof(uint64_t));
    g->fids[l] = (uint64_t *)malloc(g->F[l] * sizeof(uint64_t));
    uint64_t k = 0;
    for (uint64_t j = 0; j < n; ++j)
      if (l == N - 1 || j == 0 || fd[j] <= l) {
        start[l][k] = j;
        g->fids[l][k] = key[l][perm[j]];
        ++k;
      }
    start[l][g->F[l]] = n;
  }
  for (uint32_t l = 0; l + 1 < N; ++l) {
    g->fptr[l] = (uint64_t *)malloc((g->F[l] + 1) * sizeof(uint64_t));
    uint64_t c = 0;
    for (uint64_t p = 0; p < g->F[l]; ++p) {
      while (start[l + 1][c] < start[l][p]) ++c;
      g->fptr[l][p] = c;
    }
    g->fptr[l][g->F[l]] = g->F[l + 1];
  }
  for (uint32_t l = 0; l < N; ++l) free(start[l]);
  free(fd);
  free(perm);
}

void *orc_build(uint32_t N, uint64_t nnz, uint32_t const *const *ind,
                uint64_t const *dims, double const *vals, uint32_t ngrp,
                uint32_t const *desc, uint32_t R)
{
  if (N < 3 || N > OMAXN || ngrp < 1 || ngrp > N || nnz == 0 || R == 0 ||
      nnz >= (1ull << 32))
    return NULL;
  ohandle *h = (ohandle *)calloc(1, sizeof(ohandle));
  h->N = N;
  h->R = R;
  h->ngrp = ngrp;
  h->nnz = nnz;
  memcpy(h->dims, dims, N * sizeof(uint64_t));
  for (int i = 0; i < NLOCKS; ++i) omp_init_lock(&h->locks[i]);
  uint64_t maxF = 0;
  for (uint32_t gi = 0; gi < ngrp; ++gi) {
    ogroup *g = h->g + gi;
    uint32_t const *d = desc + gi * (3 + N);
    g->memo = d[0];
    g->b = d[1];
    g->nim = d[2];
    for (uint32_t l = 0; l < N; ++l) {
      g->order[l] = d[3 + l];
      h->level_of[gi][g->order[l]] = l;
    }
    build_group(h, g, ind, vals);
    for (uint32_t l = 1; l + 1 < N; ++l) {
      if (g->F[l] > maxF) maxF = g->F[l];
      if (g->nim > 0 && l + g->nim >= g->b && l + 1 <= g->b)
        g->memo_vals[l] = (double *)calloc(g->F[l] * R, sizeof(double));
    }
  }
  for (uint32_t gi = 0; gi < ngrp; ++gi) {
    uint32_t const m0 = h->g[gi].memo;
    uint32_t const m1 = (gi + 1 < ngrp) ? h->g[gi + 1].memo : N;
    h->use_grp[m0] = gi;
    h->use_tag[m0] = 0;
    for (uint32_t m = m0 + 1; m < m1; ++m) {
      h->use_grp[m] = gi;
      h->use_tag[m] = 1;
    }
  }
  h->tmp[0] = (double *)malloc((maxF ? maxF : 1) * R * sizeof(double));
  h->tmp[1] = (double *)malloc((maxF ? maxF : 1) * R * sizeof(double));
  double ss = 0.0;
  for (uint64_t j = 0; j < nnz; ++j) ss += vals[j] * vals[j];
  h->ttnormsq = ss;
  return h;
}

void orc_free(void *vh)
{
  ohandle *h = (ohandle *)vh;
  if (!h) return;
  for (uint32_t gi = 0; gi < h->ngrp; ++gi) {
    ogroup *g = h->g + gi;
    for (uint32_t l = 0; l < OMAXN; ++l) {
      free(g->fptr[l]);
      free(g->fids[l]);
      free(g->memo_vals[l]);
    }
    free(g->vals);
  }
  free(h->tmp[0]);
  free(h->tmp[1]);
  for (int i = 0; i < NLOCKS; ++i) omp_destroy_lock(&h->locks[i]);
  free(h);
}

int orc_get_nfibs(void *vh, uint32_t gi, uint64_t *F)
{
  ohandle *h = (ohandle *)vh;
  memcpy(F, h->g[gi].F, h->N * sizeof(uint64_t));
  return 0;
}

/* ---- level passes (p_spttm / p_spt2t_add_hada / p_spt2m_add_hada) -------- */
static void level_pass(ohandle const *h, ogroup const *g, uint32_t l,
                       double const *in, int leaf, double const *U,
                       double *out)
{
  uint32_t const R = h->R;
  uint64_t const np = g->F[l - 1];
  uint64_t const *fp = g->fptr[l - 1];
  uint64_t const *id = g->fids[l];
  #pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < (int64_t)np; ++p) {
    double *o = out + (uint64_t)p * R;
    for (uint32_t r = 0; r < R; ++r) o[r] = 0.0;
    for (uint64_t c = fp[p]; c < fp[p + 1]; ++c) {
      double const *u = U + id[c] * R;
      if (leaf) {
        double const x = in[c];
        for (uint32_t r = 0; r < R; ++r) o[r] += x * u[r];
      } else {
        double const *z = in + c * R;
        for (uint32_t r = 0; r < R; ++r) o[r] += z[r] * u[r];
      }
    }
  }
}

static int is_memo(ogroup const *g, uint32_t l)
{
  return g->nim > 0 && l + g->nim >= g->b && l + 1 <= g->b;
}

static void memo_root(ohandle *h, ogroup *g, double *const *U, double *M)
{
  uint32_t const N = h->N;
  double const *in = g->vals;
  int leaf = 1, pp = 0;
  for (uint32_t l = N - 1; l >= 1; --l) {
    double *dst;
    if (l == 1)
      dst = M;
    else if (is_memo(g, l - 1))
      dst = g->memo_vals[l - 1];
    else {
      dst = h->tmp[pp];
      pp ^= 1;
    }
    level_pass(h, g, l, in, leaf, U[g->order[l]], dst);
    in = dst;
    leaf = 0;
  }
}

/* ---- depth-first Z_l(node) (p_propagate_up / p_propagate_up_reuse) ------- */
/* Z_l(node) into bufs[l]: from the memo at level L, or from the leaves when
 * L == N-1 (then Z_{N-2}(p) = sum_c x_c U_{N-1}[id_c]). */
static void zsub(ohandle const *h, ogroup const *g, uint32_t l, uint64_t node,
                 uint32_t L, double const *Z, double *const *U, double *bufs)
{
  uint32_t const N = h->N, R = h->R;
  double *out = bufs + (size_t)l * R;
  if (l == L) {
    for (uint32_t r = 0; r < R; ++r) out[r] = Z[node * R + r];
    return;
  }
  for (uint32_t r = 0; r < R; ++r) out[r] = 0.0;
  uint64_t const c0 = g->fptr[l][node], c1 = g->fptr[l][node + 1];
  double const *Ul = U[g->order[l + 1]];
  if (l + 1 == N - 1) {
    for (uint64_t c = c0; c < c1; ++c) {
      double const x = g->vals[c];
      double const *u = Ul + g->fids[l + 1][c] * R;
      for (uint32_t r = 0; r < R; ++r) out[r] += x * u[r];
    }
    return;
  }
  double const *child = bufs + (size_t)(l + 1) * R;
  for (uint64_t c = c0; c < c1; ++c) {
    zsub(h, g, l + 1, c, L, Z, U, bufs);
    double const *u = Ul + g->fids[l + 1][c] * R;
    for (uint32_t r = 0; r < R; ++r) out[r] += child[r] * u[r];
  }
}

/* plain root: per root slice, Z_0 from the leaves (p_csf_mttkrp_root) */
static void plain_root(ohandle *h, ogroup *g, double *const *U, double *M)
{
  uint32_t const N = h->N, R = h->R;
  #pragma omp parallel
  {
    double *bufs = (double *)malloc((size_t)N * R * sizeof(double));
    #pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < (int64_t)g->F[0]; ++s) {
      zsub(h, g, 0, (uint64_t)s, N - 1, NULL, U, bufs);
      for (uint32_t r = 0; r < R; ++r) M[(uint64_t)s * R + r] = bufs[r];
    }
    free(bufs);
  }
}

/* reuse MTTKRP of level d: per root slice, depth-first down to level d
 * carrying Y = prod of ancestor factor rows; Z_d from the memo at level L
 * (propagated up when L > d); output rows under striped locks */
static void reuse_mode(ohandle *h, ogroup *g, uint32_t d, double *const *U,
                       double *M, uint64_t Mrows)
{
  uint32_t const N = h->N, R = h->R;
  uint32_t const bd = g->b - d;
  uint32_t const L = g->b - (g->nim < bd ? g->nim : bd);
  double const *Z = g->memo_vals[L];
  memset(M, 0, Mrows * R * sizeof(double));
  #pragma omp parallel
  {
    double *bufs = (double *)malloc((size_t)N * R * sizeof(double));
    double *y = (double *)malloc((size_t)(d + 1) * R * sizeof(double));
    uint64_t idx[OMAXN], end[OMAXN];
    #pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < (int64_t)g->F[0]; ++s) {
      idx[0] = (uint64_t)s;
      end[0] = (uint64_t)s + 1;
      double const *u0 = U[g->order[0]] + (uint64_t)s * R;
      for (uint32_t r = 0; r < R; ++r) y[r] = u0[r];
      uint32_t l = 1;
      idx[1] = g->fptr[0][s];
      end[1] = g->fptr[0][s + 1];
      while (l >= 1) {
        if (idx[l] == end[l]) { /* level exhausted: go up */
          --l;
          if (l >= 1) ++idx[l];
          continue;
        }
        if (l < d) { /* extend Y and descend */
          double const *ul = U[g->order[l]] + g->fids[l][idx[l]] * R;
          for (uint32_t r = 0; r < R; ++r)
            y[(size_t)l * R + r] = y[(size_t)(l - 1) * R + r] * ul[r];
          idx[l + 1] = g->fptr[l][idx[l]];
          end[l + 1] = g->fptr[l][idx[l] + 1];
          ++l;
          continue;
        }
        /* level-d node */
        uint64_t const n = idx[d];
        zsub(h, g, d, n, L, Z, U, bufs);
        double const *zd = bufs + (size_t)d * R;
        double const *yp = y + (size_t)(d - 1) * R;
        uint64_t const row = g->fids[d][n];
        double *o = M + row * R;
        omp_set_lock(&h->locks[row % NLOCKS]);
        for (uint32_t r = 0; r < R; ++r) o[r] += yp[r] * zd[r];
        omp_unset_lock(&h->locks[row % NLOCKS]);
        ++idx[d];
      }
    }
    free(bufs);
    free(y);
  }
}

int orc_mttkrp(void *vh, uint32_t mode, double *const *U, double *M)
{
  ohandle *h = (ohandle *)vh;
  if (!h || mode >= h->N) return -1;
  uint32_t const gi = h->use_grp[mode];
  ogroup *g = h->g + gi;
  if (h->use_tag[mode] == 0) {
    if (g->nim > 0)
      memo_root(h, g, U, M);
    else
      plain_root(h, g, U, M);
  } else {
    reuse_mode(h, g, h->level_of[gi][mode], U, M, h->dims[mode]);
  }
  return 0;
}

/* ---- dense helpers ---------------------------------------------------------- */
static void gram(double const *A, uint64_t I, uint32_t R, double *G)
{
  memset(G, 0, (size_t)R * R * sizeof(double));
  #pragma omp parallel
  {
    double *loc = (double *)calloc((size_t)R * R, sizeof(double));
    #pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)I; ++i)
      for (uint32_t a = 0; a < R; ++a)
        for (uint32_t b = 0; b < R; ++b)
          loc[a * R + b] += A[i * R + a] * A[i * R + b];
    #pragma omp critical
    for (uint32_t x = 0; x < R * R; ++x) G[x] += loc[x];
    free(loc);
  }
}

/* V^-1 of an SPD matrix by Cholesky (V = L L^T, then solve for each column) */
static int spd_inverse(double const *V, uint32_t R, double *Vinv)
{
  double *L = (double *)calloc((size_t)R * R, sizeof(double));
  int ok = 1;
  for (uint32_t j = 0; j < R; ++j) {
    double s = V[j * R + j];
    for (uint32_t k = 0; k < j; ++k) s -= L[j * R + k] * L[j * R + k];
    if (!(s > 0.0)) ok = 0;
    double const d = sqrt(s);
    L[j * R + j] = d;
    for (uint32_t i = j + 1; i < R; ++i) {
      double t = V[i * R + j];
      for (uint32_t k = 0; k < j; ++k) t -= L[i * R + k] * L[j * R + k];
      L[i * R + j] = t / d;
    }
  }
  double *y = (double *)malloc(R * sizeof(double));
  for (uint32_t c = 0; c < R; ++c) {
    for (uint32_t i = 0; i < R; ++i) {
      double t = (i == c) ? 1.0 : 0.0;
      for (uint32_t k = 0; k < i; ++k) t -= L[i * R + k] * y[k];
      y[i] = t / L[i * R + i];
    }
    for (int64_t i = (int64_t)R - 1; i >= 0; --i) {
      double t = y[i];
      for (uint32_t k = (uint32_t)i + 1; k < R; ++k)
        t -= L[k * R + i] * Vinv[k * R + c];
      Vinv[i * R + c] = t / L[i * R + i];
    }
  }
  free(y);
  free(L);
  return ok;
}

/* CP-ALS: U (in/out) factors, fits[niters], lambda[R]; returns sweeps run */
int orc_cpd(void *vh, double *const *U, uint32_t niters, double tol,
            double *fits, double *lambda)
{
  ohandle *h = (ohandle *)vh;
  uint32_t const N = h->N, R = h->R;
  uint64_t maxdim = 0;
  for (uint32_t m = 0; m < N; ++m)
    if (h->dims[m] > maxdim) maxdim = h->dims[m];
  double *M = (double *)malloc(maxdim * R * sizeof(double));
  double *G = (double *)malloc((size_t)N * R * R * sizeof(double));
  double *V = (double *)malloc((size_t)R * R * sizeof(double));
  double *Vi = (double *)malloc((size_t)R * R * sizeof(double));
  for (uint32_t m = 0; m < N; ++m) gram(U[m], h->dims[m], R, G + (size_t)m * R * R);
  double oldfit = 0.0;
  uint32_t it;
  for (it = 0; it < niters; ++it) {
    for (uint32_t m = 0; m < N; ++m) {
      orc_mttkrp(h, m, U, M);
      for (uint32_t x = 0; x < R * R; ++x) {
        double v = 1.0;
        for (uint32_t k = 0; k < N; ++k)
          if (k != m) v *= G[(size_t)k * R * R + x];
        V[x] = v;
      }
      spd_inverse(V, R, Vi);
      uint64_t const I = h->dims[m];
      double *A = U[m];
      #pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)I; ++i)
        for (uint32_t c = 0; c < R; ++c) {
          double s = 0.0;
          for (uint32_t j = 0; j < R; ++j) s += M[i * R + j] * Vi[j * R + c];
          A[i * R + c] = s;
        }
      for (uint32_t c = 0; c < R; ++c) {
        double v = (it == 0) ? 0.0 : -1.0e300;
        for (uint64_t i = 0; i < I; ++i) {
          double const a = A[i * R + c];
          if (it == 0)
            v += a * a;
          else
            v = a > v ? a : v;
        }
        lambda[c] = (it == 0) ? sqrt(v) : (v > 1.0 ? v : 1.0);
      }
      #pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)I; ++i)
        for (uint32_t c = 0; c < R; ++c) A[i * R + c] /= lambda[c];
      gram(A, I, R, G + (size_t)m * R * R);
    }
    double norm_mats = 0.0;
    for (uint32_t i = 0; i < R; ++i)
      for (uint32_t j = 0; j < R; ++j) {
        double hh = 1.0;
        for (uint32_t k = 0; k < N; ++k) hh *= G[(size_t)k * R * R + i * R + j];
        norm_mats += hh * lambda[i] * lambda[j];
      }
    norm_mats = fabs(norm_mats);
    double inner = 0.0;
    uint64_t const I = h->dims[N - 1];
    for (uint32_t c = 0; c < R; ++c) {
      double s = 0.0;
      for (uint64_t i = 0; i < I; ++i) s += U[N - 1][i * R + c] * M[i * R + c];
      inner += s * lambda[c];
    }
    double const fit =
        1.0 - sqrt(h->ttnormsq + norm_mats - 2.0 * inner) / sqrt(h->ttnormsq);
    fits[it] = fit;
    if (it > 0 && fabs(fit - oldfit) < tol) {
      ++it;
      break;
    }
    oldfit = fit;
  }
  free(M);
  free(G);
  free(V);
  free(Vi);
  return (int)it;
}

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int dcmp(void const *a, void const *b)
{
  double const x = *(double const *)a, y = *(double const *)b;
  return (x > y) - (x < y);
}

/* median seconds of one full MTTKRP sequence (what = 0) or one ALS sweep
 * (what = 1) after one warm-up; threads = 0 keeps the OpenMP default */
double orc_time(void *vh, double *const *U, int what, int nrep, int threads)
{
  ohandle *h = (ohandle *)vh;
  uint32_t const N = h->N, R = h->R;
  if (threads > 0) omp_set_num_threads(threads);
  uint64_t maxdim = 0;
  for (uint32_t m = 0; m < N; ++m)
    if (h->dims[m] > maxdim) maxdim = h->dims[m];
  double *M = (double *)malloc(maxdim * R * sizeof(double));
  double *t = (double *)malloc((size_t)(nrep + 1) * sizeof(double));
  double fits[2], lam[256];
  for (int rep = 0; rep <= nrep; ++rep) {
    double const t0 = now();
    if (what == 0)
      for (uint32_t m = 0; m < N; ++m) orc_mttkrp(h, m, U, M);
    else
      orc_cpd(h, U, 1, 0.0, fits, lam);
    t[rep] = now() - t0;
  }
  qsort(t + 1, (size_t)nrep, sizeof(double), dcmp);
  double const med = (nrep % 2) ? t[1 + nrep / 2]
                                : 0.5 * (t[nrep / 2] + t[1 + nrep / 2]);
  free(t);
  free(M);
  return med;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
