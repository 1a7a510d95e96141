/*
 * Replays, against include/splatt.h and libsplatt_adatm.so only, the call
 * sequences of the reference's AdaTM drivers:
 *   bench_adaptm (src/adatm_bench.c:531-650): default options, product
 *     order, decide_parameters_auto, csf_alloc_adaptive, csf/rcsf storage,
 *     rcsf_alloc_adaptive, decide_use_csfs, one mttkrp_csf_adaptive per
 *     mode, csf_free;
 *   splatt_cpd_cmd_adaptive (src/cmds/cmd_cpd.c:132-259): the same plan and
 *     structures, splatt_cpd_als_adaptive, splatt_free_kruskal,
 *     splatt_csf_free;
 *   cpd_als_iterate_adaptive (adatm_cpd.h:29-41) on caller matrices;
 *   splatt_cpd_cmd (src/cmds/cmd_cpd.c:283-312): splatt_csf_alloc +
 *     splatt_cpd_als;
 *   splatt_csf_load + splatt_mttkrp on a coordinate file.
 * Results go to binary files for the test to compare with the oracle.
 *
 * usage: boundary_replay TENSOR RANK NITERS OUTDIR LOADFILE
 *   TENSOR = a coordinate file (tt_read) or synth:CONFIG:NDRAWS (a named
 *   workload's generator, canonicalised).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "splatt.h"
#include "adatm.h"

#define FAC_SEED 7u  /* factors of the MTTKRP sequence */
#define ITER_SEED 9u /* initial factors of cpd_als_iterate_adaptive */
#define CPD_SEED 5.0 /* SPLATT_OPTION_RANDSEED of the CP-ALS drivers */

static char const *g_out;

static void die(char const *what)
{
  fprintf(stderr, "boundary_replay: %s failed: %s\n", what, adatm_last_error());
  exit(1);
}

static void put(char const *name, double const *v, size_t n)
{
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", g_out, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(v, sizeof(double), n, f) != n) die(name);
  fclose(f);
}

static sptensor_t *synth(int config, uint64_t ndraws)
{
  adatm_synth_spec spec;
  adatm_coo t;
  if (adatm_synth_named(config, ndraws, &spec) != SPLATT_SUCCESS ||
      adatm_synth(&spec, &t) != SPLATT_SUCCESS ||
      adatm_coo_canonicalize(&t, NULL) != SPLATT_SUCCESS)
    die("synth");
  sptensor_t *tt = (sptensor_t *)calloc(1, sizeof(sptensor_t));
  tt->nnz = t.nnz;
  tt->nmodes = t.nmodes;
  tt->dims = (splatt_idx_t *)malloc(t.nmodes * sizeof(splatt_idx_t));
  tt->ind = (splatt_idx_t **)malloc(t.nmodes * sizeof(splatt_idx_t *));
  tt->vals = (splatt_val_t *)malloc(t.nnz * sizeof(splatt_val_t));
  for (uint32_t m = 0; m < t.nmodes; ++m) {
    tt->dims[m] = t.dims[m];
    tt->ind[m] = (splatt_idx_t *)malloc(t.nnz * sizeof(splatt_idx_t));
    for (uint64_t j = 0; j < t.nnz; ++j) tt->ind[m][j] = t.ind[m][j];
  }
  memcpy(tt->vals, t.vals, t.nnz * sizeof(double));
  adatm_coo_free(&t);
  return tt;
}

static void put_kruskal(char const *name, splatt_kruskal const *k, splatt_idx_t nmodes,
                        splatt_idx_t const *dims)
{
  size_t n = 1 + k->rank;
  for (splatt_idx_t m = 0; m < nmodes; ++m) n += dims[m] * k->rank;
  double *buf = (double *)malloc(n * sizeof(double)), *p = buf;
  *p++ = k->fit;
  memcpy(p, k->lambda, k->rank * sizeof(double));
  p += k->rank;
  for (splatt_idx_t m = 0; m < nmodes; ++m) {
    memcpy(p, k->factors[m], dims[m] * k->rank * sizeof(double));
    p += dims[m] * k->rank;
  }
  put(name, buf, n);
  free(buf);
}

int main(int argc, char **argv)
{
  if (argc != 6) {
    fprintf(stderr, "usage: %s TENSOR RANK NITERS OUTDIR LOADFILE\n", argv[0]);
    return 2;
  }
  splatt_idx_t const nfactors = (splatt_idx_t)atoi(argv[2]);
  int const niters = atoi(argv[3]);
  g_out = argv[4];
  sptensor_t *tt;
  if (strncmp(argv[1], "synth:", 6) == 0) {
    int config = 0;
    unsigned long long nd = 0;
    if (sscanf(argv[1] + 6, "%d:%llu", &config, &nd) != 2) return 2;
    tt = synth(config, nd);
  } else {
    tt = tt_read(argv[1]);
  }
  if (!tt) die("tt_read");
  splatt_idx_t const nmodes = tt->nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  for (splatt_idx_t m = 0; m < nmodes; ++m) dims[m] = tt->dims[m];

  /* ---- bench_adaptm ---- */
  double *cpd_opts = splatt_default_opts();
  cpd_opts[SPLATT_OPTION_CSF_ALLOC] = SPLATT_ADAPTIVE_TIME_EFFICIENT;
  cpd_opts[SPLATT_OPTION_TILE] = SPLATT_NOTILE;
  cpd_opts[SPLATT_OPTION_NTHREADS] = 1;
  splatt_idx_t **product_order = (splatt_idx_t **)malloc(nmodes * sizeof(splatt_idx_t *));
  for (splatt_idx_t i = 0; i < nmodes; ++i)
    product_order[i] = (splatt_idx_t *)malloc(2 * nmodes * sizeof(splatt_idx_t));
  decide_product_order(tt, product_order);
  group_properties *grp_prop = NULL;
  splatt_idx_t const n_grp = decide_parameters_auto(
      tt, nfactors, 1, (splatt_idx_t const **)product_order, &grp_prop);
  if (n_grp < 1 || !grp_prop) die("decide_parameters_auto");
  for (splatt_idx_t i = 0; i < nmodes; ++i) free(product_order[i]);
  free(product_order);
  splatt_idx_t n_csf = 0;
  splatt_csf *cs = csf_alloc_adaptive(tt, grp_prop, n_grp, cpd_opts, &n_csf);
  if (!cs) die("csf_alloc_adaptive");
  size_t const cs_bytes = csf_storage_adaptive(cs, n_csf, grp_prop, (int)n_grp);
  splatt_idx_t n_rcsf = 0;
  rcsf_seq_adaptive *rs_seq =
      rcsf_alloc_adaptive(cs, nfactors, grp_prop, n_grp, n_csf, cpd_opts, &n_rcsf);
  if (!rs_seq) die("rcsf_alloc_adaptive");
  size_t const rcsf_bytes = rcsf_storage_adaptive(rs_seq, n_rcsf);
  splatt_idx_t use_csfs[SPLATT_MAX_NMODES], use_tags[SPLATT_MAX_NMODES];
  decide_use_csfs(nmodes, grp_prop, (int)n_grp, n_csf, use_csfs, use_tags);

  /* the plan, structure counts and storage as text */
  char path[4096];
  snprintf(path, sizeof(path), "%s/plan.txt", g_out);
  FILE *pf = fopen(path, "w");
  if (!pf) die("plan.txt");
  fprintf(pf, "n_grp %llu\nn_csf %llu\nn_rcsf %llu\ncs_bytes %zu\nrcsf_bytes %zu\n",
          (unsigned long long)n_grp, (unsigned long long)n_csf,
          (unsigned long long)n_rcsf, cs_bytes, rcsf_bytes);
  for (splatt_idx_t g = 0; g < n_grp; ++g) {
    fprintf(pf, "group %llu %llu %llu", (unsigned long long)grp_prop[g].memo_mode,
            (unsigned long long)grp_prop[g].begin_imten,
            (unsigned long long)grp_prop[g].n_imten);
    for (splatt_idx_t m = 0; m < nmodes; ++m)
      fprintf(pf, " %llu", (unsigned long long)grp_prop[g].o_opt[m]);
    fprintf(pf, "\n");
  }
  for (splatt_idx_t g = 0; g < n_csf; ++g) {
    fprintf(pf, "nfibs %llu", (unsigned long long)g);
    for (splatt_idx_t l = 0; l < nmodes; ++l)
      fprintf(pf, " %llu", (unsigned long long)cs[g].pt->nfibs[l]);
    fprintf(pf, "\n");
  }
  fprintf(pf, "use");
  for (splatt_idx_t m = 0; m < nmodes; ++m)
    fprintf(pf, " %llu:%llu", (unsigned long long)use_csfs[m],
            (unsigned long long)use_tags[m]);
  fprintf(pf, "\n");
  fclose(pf);

  matrix_t *mats[SPLATT_MAX_NMODES + 1];
  splatt_idx_t maxdim = 0;
  for (splatt_idx_t m = 0; m < nmodes; ++m) {
    mats[m] = mat_alloc(dims[m], nfactors);
    adatm_rand_matrix(FAC_SEED, (uint32_t)m, dims[m], nfactors, mats[m]->vals);
    if (dims[m] > maxdim) maxdim = dims[m];
  }
  mats[SPLATT_MAX_NMODES] = mat_alloc(maxdim, nfactors);
  for (splatt_idx_t m = 0; m < nmodes; ++m) {
    mttkrp_csf_adaptive(cs, rs_seq, n_csf, mats, m, NULL, grp_prop, (int)n_grp,
                        use_csfs[m], use_tags[m], cpd_opts);
    char name[64];
    snprintf(name, sizeof(name), "mttkrp_%llu.bin", (unsigned long long)m);
    put(name, mats[SPLATT_MAX_NMODES]->vals, dims[m] * nfactors);
  }

  /* ---- splatt_cpd_cmd_adaptive ---- */
  double *opts = splatt_default_opts();
  opts[SPLATT_OPTION_CSF_ALLOC] = SPLATT_ADAPTIVE_TIME_EFFICIENT;
  opts[SPLATT_OPTION_NITER] = niters;
  opts[SPLATT_OPTION_TOLERANCE] = 0.0;
  opts[SPLATT_OPTION_RANDSEED] = CPD_SEED;
  opts[SPLATT_OPTION_VERBOSITY] = SPLATT_VERBOSITY_NONE;
  splatt_kruskal factored;
  if (splatt_cpd_als_adaptive(cs, rs_seq, n_csf, grp_prop, n_grp, use_csfs, use_tags,
                              nfactors, opts, &factored) != SPLATT_SUCCESS)
    die("splatt_cpd_als_adaptive");
  put_kruskal("cpd_adaptive.bin", &factored, nmodes, dims);
  splatt_free_kruskal(&factored);

  /* ---- cpd_als_iterate_adaptive on caller matrices ---- */
  for (splatt_idx_t m = 0; m < nmodes; ++m)
    adatm_rand_matrix(ITER_SEED, (uint32_t)m, dims[m], nfactors, mats[m]->vals);
  splatt_val_t *lambda = (splatt_val_t *)malloc(nfactors * sizeof(splatt_val_t));
  double const fit = cpd_als_iterate_adaptive(cs, rs_seq, n_csf, grp_prop, n_grp, use_csfs,
                                              use_tags, mats, lambda, nfactors, NULL, opts);
  if (fit != fit) die("cpd_als_iterate_adaptive");
  {
    splatt_kruskal k;
    memset(&k, 0, sizeof(k));
    k.rank = nfactors;
    k.fit = fit;
    k.lambda = lambda;
    for (splatt_idx_t m = 0; m < nmodes; ++m) k.factors[m] = mats[m]->vals;
    put_kruskal("iterate.bin", &k, nmodes, dims);
  }
  free(lambda);
  for (splatt_idx_t m = 0; m < nmodes; ++m) mat_free(mats[m]);
  mat_free(mats[SPLATT_MAX_NMODES]);
  csf_free(cs, cpd_opts); /* bench_adaptm's release (rs_seq's structure aliases cs) */
  for (splatt_idx_t j = 0; j < n_rcsf; ++j) {
    for (splatt_idx_t k = 0; k < rs_seq[j].n_imten; ++k) free(rs_seq[j].rcsfs[k].pt);
    free(rs_seq[j].rcsfs);
  }
  free(rs_seq);
  for (splatt_idx_t g = 0; g < n_grp; ++g) free(grp_prop[g].o_opt);
  free(grp_prop);
  free(cpd_opts);

  /* ---- splatt_cpd_cmd: splatt_csf_alloc + splatt_cpd_als ---- */
  splatt_csf *csf = splatt_csf_alloc(tt, opts);
  if (!csf) die("splatt_csf_alloc");
  tt_free(tt);
  if (splatt_cpd_als(csf, nfactors, opts, &factored) != SPLATT_SUCCESS) die("splatt_cpd_als");
  put_kruskal("cpd_plain.bin", &factored, nmodes, dims);
  splatt_free_kruskal(&factored);
  splatt_csf_free(csf, opts);

  /* ---- splatt_csf_load + splatt_mttkrp ---- */
  splatt_idx_t ln = 0;
  splatt_csf *lc = NULL;
  if (splatt_csf_load(argv[5], &ln, &lc, opts) != SPLATT_SUCCESS) die("splatt_csf_load");
  splatt_val_t *U[SPLATT_MAX_NMODES];
  splatt_idx_t lmax = 0;
  for (splatt_idx_t m = 0; m < ln; ++m) {
    U[m] = (splatt_val_t *)malloc(lc->dims[m] * nfactors * sizeof(splatt_val_t));
    adatm_rand_matrix(FAC_SEED, (uint32_t)m, lc->dims[m], nfactors, U[m]);
    if (lc->dims[m] > lmax) lmax = lc->dims[m];
  }
  splatt_val_t *out = (splatt_val_t *)malloc(lmax * nfactors * sizeof(splatt_val_t));
  for (splatt_idx_t m = 0; m < ln; ++m) {
    if (splatt_mttkrp(m, nfactors, lc, U, out, opts) != SPLATT_SUCCESS) die("splatt_mttkrp");
    char name[64];
    snprintf(name, sizeof(name), "load_mttkrp_%llu.bin", (unsigned long long)m);
    put(name, out, lc->dims[m] * nfactors);
  }
  for (splatt_idx_t m = 0; m < ln; ++m) free(U[m]);
  free(out);
  splatt_free_csf(lc, opts);
  splatt_free_opts(opts);
  printf("boundary_replay ok\n");
  return 0;
}
