/*
 * `splatt` command-line tool: the AdaTM subcommands on one MI355X.
 *
 *   splatt bench TENSOR -a adatm|csf [-a ...] [-i NITERS] [-r RANK] [-t NTHREADS]
 *                [-s] [-w] [--tile] [--precision P] [--planner PL]
 *   splatt cpd   TENSOR [-i NITERS] [--tol TOL] [-r RANK] [-t NTHREADS] [--tile]
 *                [--nowrite] [-v] [--precision P] [--seed S] [--planner PL]
 *
 * Flags, defaults and log lines follow the reference's drivers:
 * src/cmds/cmd_bench.c:75-87, 203-287 with src/adatm_bench.c:516-655
 * (bench: -i 5, -r 16, -t 1; prints the plan, CSF-/vCSF-STORAGE, per-mode and
 * per-iteration times) and src/cmds/cmd_cpd.c:23-32, 123-261 with
 * src/adatm_cpd.c:256-374 (cpd: -i 50, --tol 1e-5, -r 10; prints the fit per
 * iteration, "Final fit", writes lambda.mat and mode<m>.mat).
 * TENSOR is a coordinate text file (.tns: N 1-based indices and a value per
 * line) or synth:<1|2|3>[:draws] for the seeded synthetic workloads.
 * Duplicate coordinates are summed and empty indices removed on input.
 * Reordering follows cmd_bench.c:119-149, 243-251: -p FILE relabels the
 * tensor with the partition FILE by the type of -z (graph, hgraph, fib;
 * basis mode -m, 1-based, default 1); -z without -p does nothing; -p
 * without a known -z is an error.  The factor matrices are drawn in the
 * original labels and follow their rows to the new ones
 * (adatm_bench.c:40-51, 529), and -w writes every MTTKRP back in the
 * original labels, so the outputs do not depend on the reordering.
 * --tile gives -a csf SPLATT's dense tiling (SPLATT_DENSETILE, one tile
 * per thread along every level below the root); the AdaTM CSFs are
 * untiled as in the reference (adatm_bench.c:535).  The options of this build
 * (--precision, --planner, --seed) are long options only, so every short
 * option keeps the reference's meaning.
 */
#include <argp.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../../include/adatm.h"
#include "../../../include/splatt.h"

int adatm_read_tns(char const *fname, uint32_t *nmodes_out, uint64_t *nnz_out,
                   splatt_idx_t ***inds_out, splatt_val_t **vals_out);

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int load_tensor(char const *name, adatm_coo *t)
{
  memset(t, 0, sizeof(*t));
  if (strncmp(name, "synth:", 6) == 0) {
    int cfg = atoi(name + 6);
    char const *c = strchr(name + 6, ':');
    uint64_t draws = c ? strtoull(c + 1, NULL, 10) : 0;
    adatm_synth_spec spec;
    if (adatm_synth_named(cfg, draws, &spec) != SPLATT_SUCCESS) return -1;
    if (adatm_synth(&spec, t) != SPLATT_SUCCESS) return -1;
  } else {
    uint32_t N;
    uint64_t nnz;
    splatt_idx_t **ind;
    splatt_val_t *val;
    if (adatm_read_tns(name, &N, &nnz, &ind, &val) != SPLATT_SUCCESS) return -1;
    if (N < 3 || N > ADATM_MAX_NMODES) return -1;
    uint64_t dims[ADATM_MAX_NMODES];
    for (uint32_t m = 0; m < N; ++m) {
      uint64_t mx = 0;
      for (uint64_t j = 0; j < nnz; ++j) mx = ind[m][j] > mx ? ind[m][j] : mx;
      dims[m] = mx + 1;
    }
    if (adatm_coo_alloc(N, nnz, dims, t) != SPLATT_SUCCESS) return -1;
    for (uint32_t m = 0; m < N; ++m) {
      for (uint64_t j = 0; j < nnz; ++j) t->ind[m][j] = (uint32_t)ind[m][j];
      free(ind[m]);
    }
    memcpy(t->vals, val, nnz * sizeof(double));
    free(ind);
    free(val);
  }
  return adatm_coo_canonicalize(t, NULL) == SPLATT_SUCCESS ? 0 : -1;
}

static char const *bytes_str(double b, char *buf)
{
  char const *u[] = {"B", "KB", "MB", "GB", "TB"};
  int k = 0;
  while (b >= 1024.0 && k < 4) {
    b /= 1024.0;
    ++k;
  }
  sprintf(buf, "%0.2f%s", b, u[k]);
  return buf;
}

static void print_plan(adatm_plan const *p, uint32_t N)
{
  printf("====== Optimal grp_prop =======\n");
  printf("[memo_mode, begin_imten, n_imten, o_opt]\n");
  for (uint32_t g = 0; g < p->n_grp; ++g) {
    printf("%u, %u, %u, \n", p->grp[g].memo_mode, p->grp[g].begin_imten,
           p->grp[g].n_imten);
    printf("o_opt: ");
    for (uint32_t m = 0; m < N; ++m) printf("%u ", p->grp[g].o_opt[m]);
    printf("\n");
  }
  printf("\n");
}

static void print_header(adatm_coo const *t, char const *name)
{
  printf("****************************************************************\n");
  printf("splatt (MI355X AdaTM) v%d.%d.%d\n\n", SPLATT_VER_MAJOR,
         SPLATT_VER_MINOR, SPLATT_VER_SUBMINOR);
  printf("Tensor information ---------------------------------------------\n");
  printf("FILE=%s\nDIMS=", name);
  for (uint32_t m = 0; m < t->nmodes; ++m)
    printf("%llu%s", (unsigned long long)t->dims[m], m + 1 < t->nmodes ? "x" : "");
  printf(" NNZ=%llu\n\n", (unsigned long long)t->nnz);
}

/* planner: 0 analytic (the reference's), 1 learned (cost model beside the
 * library or $ADATM_COSTMODEL; analytic without one), 2 exhaustive timing */
static int plan_for(adatm_coo const *t, uint64_t R, int planner, int prec,
                    adatm_plan *plan)
{
  adatm_plan cands[256];
  int nc = 0, ch = 0, rc;
  if (planner == SPLATT_ADATM_PLAN_LEARNED) {
    adatm_mlp *model = NULL;
    char path[4096];
    if (adatm_costmodel_path(path, sizeof(path)) &&
        adatm_mlp_load(0, path, &model) != SPLATT_SUCCESS)
      model = NULL;
    printf("Planner: %s\n", model ? "learned cost model" : "analytic (no cost model)");
    rc = adatm_plan_learned(t, R, model, 0, cands, 256, &nc, &ch, NULL);
    adatm_mlp_destroy(model);
  } else if (planner == SPLATT_ADATM_PLAN_EXHAUSTIVE) {
    printf("Planner: exhaustive timing\n");
    rc = adatm_plan_exhaustive(t, R, prec, 0, 1, 3, cands, 256, &nc, &ch, NULL);
  } else {
    rc = adatm_plan_analytic(t, R, 1, (double)SPLATT_ADATM_SLIMIT, 0, cands, 256, &nc,
                             &ch);
  }
  if (rc != SPLATT_SUCCESS) return -1;
  printf("Number of configurations: %d\n", nc);
  char b[32];
  printf("Storage limitation: %s\n\n", bytes_str((double)SPLATT_ADATM_SLIMIT, b));
  *plan = cands[ch];
  printf("Optimal predicted group properties:\nNumber of groups: %u\n",
         plan->n_grp);
  print_plan(plan, t->nmodes);
  return 0;
}

static void print_storage(adatm_engine *e, adatm_plan const *plan, uint32_t N)
{
  uint64_t F[ADATM_MAX_NMODES * ADATM_MAX_NMODES];
  double sb, sf;
  uint64_t db;
  adatm_engine_stats(e, F, &sb, &sf, &db);
  uint64_t memo = 0;
  for (uint32_t g = 0; g < plan->n_grp; ++g)
    for (uint32_t l = 0; l < N; ++l) memo += adatm_engine_memo_bytes(e, g, l);
  char b1[32], b2[32];
  printf("Number of CSFs: %u\n** CSF **\nCSF-STORAGE: %s\n\n", plan->n_grp,
         bytes_str((double)(db - memo), b1));
  printf("** vCSF **\nvCSF-STORAGE: %s\n\n\n", bytes_str((double)memo, b2));
}

/* ---- bench ------------------------------------------------------------ */
struct bench_args {
  char *tensor;
  int adatm, csf;
  char *badalg;
  uint32_t niters, nthreads, rank;
  int prec, planner, scale, write, tile;
  int rtype;      /* splatt_perm_type of -z; PERM_ERROR when absent/unknown */
  char *permerr;  /* the unknown -z argument */
  char *pfile;    /* -p */
  uint32_t permmode;
  uint32_t *perm[ADATM_MAX_NMODES]; /* old -> new labels after -p, else NULL */
};

#define OPT_TILE 255
#define OPT_PREC 251
#define OPT_PLANNER 250
static struct argp_option bench_opts[] = {
    {"alg", 'a', "ALG", 0, "algorithm to benchmark: adatm, csf (repeatable)", 0},
    {"iters", 'i', "NITERS", 0, "number of iterations to use (default: 5)", 0},
    {"mode", 'm', "MODE", 0, "mode basis for hgraph reordering (default: 1)", 0},
    {"threads", 't', "NTHREADS", 0, "number of threads to use (default: 1)", 0},
    {"rank", 'r', "RANK", 0, "rank of decomposition to find (default: 16)", 0},
    {"scale", 's', 0, 0, "scale threads from 1 to NTHREADS (by 2)", 0},
    {"tile", OPT_TILE, 0, 0, "use tiling during SPLATT", 0},
    {"write", 'w', 0, 0, "write results to files ALG_mode<N>.mat (for testing)", 0},
    {"rtype", 'z', "TYPE", 0, "designate reordering type (graph, hgraph, fib)", 0},
    {"pfile", 'p', "FILE", 0, "partition file for reordering", 0},
    {"precision", OPT_PREC, "P", 0, "fp64 | fp32 | bf16rows (default: fp64)", 0},
    {"planner", OPT_PLANNER, "PLANNER", 0,
     "analytic | learned | exhaustive (default: analytic)", 0},
    {0}};

static int parse_planner(char const *s)
{
  if (!strcmp(s, "learned")) return SPLATT_ADATM_PLAN_LEARNED;
  if (!strcmp(s, "exhaustive")) return SPLATT_ADATM_PLAN_EXHAUSTIVE;
  return SPLATT_ADATM_PLAN_ANALYTIC;
}

static int parse_prec(char const *s)
{
  if (!strcmp(s, "fp32")) return SPLATT_ADATM_FP32;
  if (!strcmp(s, "bf16rows")) return SPLATT_ADATM_BF16ROWS;
  return SPLATT_ADATM_FP64;
}

static error_t bench_parse(int key, char *arg, struct argp_state *st)
{
  struct bench_args *a = (struct bench_args *)st->input;
  switch (key) {
  case 'a':
    if (!strcmp(arg, "adatm"))
      a->adatm = 1;
    else if (!strcmp(arg, "csf"))
      a->csf = 1;
    else
      a->badalg = arg;
    break;
  case 'i': a->niters = (uint32_t)atoi(arg); break;
  case 't': a->nthreads = (uint32_t)atoi(arg); break;
  case 'r': a->rank = (uint32_t)atoi(arg); break;
  case 's': a->scale = 1; break;
  case 'w': a->write = 1; break;
  case OPT_TILE: a->tile = 1; break;
  case 'z':
    if (!strcmp(arg, "graph"))
      a->rtype = PERM_GRAPH;
    else if (!strcmp(arg, "hgraph"))
      a->rtype = PERM_HGRAPH;
    else if (!strcmp(arg, "fib"))
      a->rtype = PERM_FIBSCHED;
    else {
      a->rtype = PERM_ERROR;
      a->permerr = arg;
    }
    break;
  case 'p': a->pfile = arg; break;
  case 'm': a->permmode = (uint32_t)(atoi(arg) - 1); break;
  case OPT_PREC: a->prec = parse_prec(arg); break;
  case OPT_PLANNER: a->planner = parse_planner(arg); break;
  case ARGP_KEY_ARG:
    if (a->tensor) argp_usage(st);
    a->tensor = arg;
    break;
  case ARGP_KEY_END:
    if (!a->tensor) argp_usage(st);
    break;
  default: return ARGP_ERR_UNKNOWN;
  }
  return 0;
}

static void write_matrix(char const *fname, double const *A, uint64_t I,
                         uint64_t J);

/* rows of a relabelled matrix back to the original labels: row i of the
 * result is row perm[i] of A (in place through a copy) */
static void unshuffle_rows(double *A, uint32_t const *perm, uint64_t I, uint64_t J)
{
  double *c = (double *)malloc(I * J * sizeof(double));
  if (!c) return;
  memcpy(c, A, I * J * sizeof(double));
  for (uint64_t i = 0; i < I; ++i) memcpy(A + i * J, c + (uint64_t)perm[i] * J, J * sizeof(double));
  free(c);
}

/* the initial factor of mode m (seed 1), drawn in the original labels and
 * moved to the new ones: row perm[i] = drawn row i (adatm_bench.c:40-51) */
static void bench_factor(struct bench_args const *a, adatm_coo const *t, uint32_t m,
                         double *A)
{
  adatm_rand_matrix(1, m, t->dims[m], a->rank, A);
  if (!a->perm[m]) return;
  double *c = (double *)malloc(t->dims[m] * a->rank * sizeof(double));
  if (!c) return;
  memcpy(c, A, t->dims[m] * a->rank * sizeof(double));
  for (uint64_t i = 0; i < t->dims[m]; ++i)
    memcpy(A + (uint64_t)a->perm[m][i] * a->rank, c + i * a->rank, a->rank * sizeof(double));
  free(c);
}

/* thread counts of a run: NTHREADS, or 1, 2, 4, ..., NTHREADS with -s
 * (cmd_bench.c:174-201); the GPU path does not use host threads, so only
 * the log lines differ */
static uint32_t thread_runs(struct bench_args const *a, uint32_t *th)
{
  uint32_t n = 0;
  if (a->scale) {
    for (uint32_t t = 1; t < a->nthreads; t *= 2) th[n++] = t;
  }
  th[n++] = a->nthreads ? a->nthreads : 1;
  return n;
}

/* the timed loop of bench_adaptm (adatm_bench.c:606-642): iterations x modes,
 * with -w the first iteration's MTTKRPs written to <alg>_mode<m>.mat */
typedef int (*mttkrp_fn)(void *ctx, uint32_t m, double *out);

static int bench_loop(struct bench_args const *a, adatm_coo const *t, char const *alg,
                      mttkrp_fn f, void *ctx)
{
  uint32_t th[40];
  uint32_t const nruns = thread_runs(a, th);
  uint64_t maxd = 0;
  for (uint32_t m = 0; m < t->nmodes; ++m) maxd = t->dims[m] > maxd ? t->dims[m] : maxd;
  double *out = a->write ? (double *)malloc(maxd * a->rank * sizeof(double)) : NULL;
  for (uint32_t r = 0; r < nruns; ++r) {
    if (nruns > 1) printf("## THREADS %u\n", th[r]);
    for (uint32_t i = 0; i < a->niters; ++i) {
      double const t0 = now();
      for (uint32_t m = 0; m < t->nmodes; ++m) {
        int const w = out && r + 1 == nruns && i == 0;
        double const tm = now();
        if (f(ctx, m, w ? out : NULL) != SPLATT_SUCCESS) {
          fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
          free(out);
          return -1;
        }
        printf("  mode %u %0.3fs\n", m + 1, now() - tm);
        if (w) {
          char fn[64];
          snprintf(fn, sizeof(fn), "%s_mode%u.mat", alg, m + 1);
          if (a->perm[m]) unshuffle_rows(out, a->perm[m], t->dims[m], a->rank);
          write_matrix(fn, out, t->dims[m], a->rank);
        }
      }
      printf("    its = %3u (%0.3fs)\n", i + 1, now() - t0);
    }
  }
  free(out);
  return 0;
}

static int adatm_mttkrp_fn(void *ctx, uint32_t m, double *out)
{
  return adatm_engine_mttkrp((adatm_engine *)ctx, m, out);
}

/* -a csf: SPLATT's plain CSF MTTKRP (two CSFs, SPLATT's default
 * allocation) through the public API */
typedef struct {
  splatt_csf *csf;
  double *opts;
  double **U;
  double *scratch;
  uint64_t rank;
} csf_ctx;

static int csf_mttkrp_fn(void *vctx, uint32_t m, double *out)
{
  csf_ctx *c = (csf_ctx *)vctx;
  return splatt_mttkrp(m, c->rank, c->csf, c->U, out ? out : c->scratch, c->opts);
}

static int bench_csf(struct bench_args const *a, adatm_coo const *t)
{
  printf("## CSF\n");
  csf_ctx c;
  memset(&c, 0, sizeof(c));
  c.rank = a->rank;
  c.opts = splatt_default_opts();
  c.opts[SPLATT_OPTION_CSF_ALLOC] = SPLATT_CSF_TWOMODE;
  c.opts[SPLATT_OPTION_ADATM_PRECISION] = a->prec;
  c.opts[SPLATT_OPTION_NTHREADS] = a->nthreads ? a->nthreads : 1;
  if (a->tile) c.opts[SPLATT_OPTION_TILE] = SPLATT_DENSETILE;
  splatt_idx_t *ind[ADATM_MAX_NMODES];
  uint64_t maxd = 0;
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    ind[m] = (splatt_idx_t *)malloc(t->nnz * sizeof(splatt_idx_t));
    for (uint64_t j = 0; j < t->nnz; ++j) ind[m][j] = t->ind[m][j];
    maxd = t->dims[m] > maxd ? t->dims[m] : maxd;
  }
  int rc = splatt_csf_convert(t->nmodes, t->nnz, ind, t->vals, &c.csf, c.opts);
  for (uint32_t m = 0; m < t->nmodes; ++m) free(ind[m]);
  if (rc != SPLATT_SUCCESS) {
    fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
    free(c.opts);
    return -1;
  }
  c.U = (double **)calloc(t->nmodes, sizeof(double *));
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    c.U[m] = (double *)malloc(t->dims[m] * a->rank * sizeof(double));
    bench_factor(a, t, m, c.U[m]);
  }
  c.scratch = (double *)malloc(maxd * a->rank * sizeof(double));
  rc = bench_loop(a, t, "csf", csf_mttkrp_fn, &c);
  for (uint32_t m = 0; m < t->nmodes; ++m) free(c.U[m]);
  free(c.U);
  free(c.scratch);
  splatt_free_csf(c.csf, c.opts);
  free(c.opts);
  return rc;
}

/* -p FILE: relabel the tensor (cmd_bench.c:237-251); a->perm gets the
 * labels so the factors and -w outputs can follow them */
static int reorder(struct bench_args *a, adatm_coo *t)
{
  if (a->rtype == PERM_ERROR) {
    fprintf(stderr, "SPLATT: reordering algorithm '%s' is not recognized.\n"
                    "Run with '--help' for assistance.\n",
            a->permerr ? a->permerr : "(null)");
    return -1;
  }
  printf("Reordering ------------------------------------------------------\n");
  double const t0 = now();
  uint64_t np = 0;
  uint32_t *parts = NULL;
  int rc = adatm_part_read(a->pfile, &np, &parts);
  for (uint32_t m = 0; rc == SPLATT_SUCCESS && m < t->nmodes; ++m)
    if (!(a->perm[m] = (uint32_t *)malloc(t->dims[m] * sizeof(uint32_t))))
      rc = SPLATT_ERROR_NOMEMORY;
  if (rc == SPLATT_SUCCESS)
    rc = adatm_perm_compute(t, a->rtype, a->permmode, parts, np, 0, a->perm);
  free(parts);
  if (rc != SPLATT_SUCCESS) {
    fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
    return -1;
  }
  adatm_coo_relabel(t, a->perm);
  static char const *const names[] = {"graph", "hgraph", "fib"};
  printf("TYPE=%s", names[a->rtype]);
  if (a->rtype != PERM_GRAPH) printf(" MODE=%u", a->permmode + 1);
  printf(" TIME=%0.3fs\n\n", now() - t0);
  return 0;
}

static int cmd_bench(int argc, char **argv)
{
  struct bench_args a;
  memset(&a, 0, sizeof(a));
  a.niters = 5;
  a.nthreads = 1;
  a.rank = 16;
  a.prec = SPLATT_ADATM_FP64;
  a.planner = SPLATT_ADATM_PLAN_ANALYTIC;
  a.rtype = PERM_ERROR;
  struct argp ap = {bench_opts, bench_parse, "TENSOR -a ALG",
                    "splatt-bench -- benchmark the memoized MTTKRP sequence.", 0, 0, 0};
  argp_parse(&ap, argc, argv, ARGP_IN_ORDER, 0, &a);
  if (a.badalg || (!a.adatm && !a.csf)) {
    fprintf(stderr, "SPLATT: algorithm '%s' is not recognized (adatm, csf).\n",
            a.badalg ? a.badalg : "");
    return EXIT_FAILURE;
  }
  adatm_coo t;
  if (load_tensor(a.tensor, &t) != 0) {
    fprintf(stderr, "SPLATT: cannot load '%s': %s\n", a.tensor, adatm_last_error());
    return EXIT_FAILURE;
  }
  print_header(&t, a.tensor);
  if (a.pfile && reorder(&a, &t) != 0) {
    adatm_coo_free(&t);
    return EXIT_FAILURE;
  }
  printf("Benchmarking ---------------------------------------------------\n");
  printf("RANK=%u ITS=%u\n", a.rank, a.niters);
  int rc = 0;
  if (a.csf) rc = bench_csf(&a, &t);
  if (rc == 0 && a.adatm) {
    printf("## ADATM\n");
    adatm_plan plan;
    if (plan_for(&t, a.rank, a.planner, a.prec, &plan) != 0) return EXIT_FAILURE;
    adatm_engine *e = NULL;
    if (adatm_engine_create(&t, a.rank, &plan, a.prec, 0, &e) != SPLATT_SUCCESS) {
      fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
      return EXIT_FAILURE;
    }
    print_storage(e, &plan, t.nmodes);
    double *A = NULL;
    uint64_t maxd = 0;
    for (uint32_t m = 0; m < t.nmodes; ++m) maxd = t.dims[m] > maxd ? t.dims[m] : maxd;
    A = (double *)malloc(maxd * a.rank * sizeof(double));
    for (uint32_t m = 0; rc == 0 && m < t.nmodes; ++m) {
      bench_factor(&a, &t, m, A);
      if (!A || adatm_engine_set_factor(e, m, A) != SPLATT_SUCCESS) rc = -1;
    }
    free(A);
    if (rc == 0) rc = bench_loop(&a, &t, "adatm", adatm_mttkrp_fn, e);
    else fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
    adatm_engine_destroy(e);
  }
  for (uint32_t m = 0; m < ADATM_MAX_NMODES; ++m) free(a.perm[m]);
  adatm_coo_free(&t);
  return rc == 0 ? EXIT_SUCCESS : EXIT_FAILURE;
}

/* ---- cpd -------------------------------------------------------------- */
struct cpd_args {
  char *tensor;
  uint32_t niters, rank, nthreads;
  double tol;
  int write, verbose, prec;
  uint64_t seed;
  int planner;
};

#define OPT_TOL 254
#define OPT_NOWRITE 253
#define OPT_SEED 252
static struct argp_option cpd_opts[] = {
    {"iters", 'i', "NITERS", 0, "maximum number of iterations to use (default: 50)", 0},
    {"tol", OPT_TOL, "TOLERANCE", 0, "minimum change for convergence (default: 1e-5)", 0},
    {"rank", 'r', "RANK", 0, "rank of decomposition to find (default: 10)", 0},
    {"threads", 't', "NTHREADS", 0, "number of threads to use (default: #cores)", 0},
    {"tile", OPT_TILE, 0, 0, "use tiling during SPLATT (AdaTM's CSFs are untiled)", 0},
    {"nowrite", OPT_NOWRITE, 0, 0, "do not write output to file", 0},
    {"verbose", 'v', 0, 0, "turn on verbose output (default: no)", 0},
    {"precision", OPT_PREC, "P", 0, "fp64 | fp32 | bf16rows (default: fp64)", 0},
    {"seed", OPT_SEED, "SEED", 0, "factor initialisation seed (default: time)", 0},
    {"planner", OPT_PLANNER, "PLANNER", 0,
     "analytic | learned | exhaustive (default: analytic)", 0},
    {0}};

static error_t cpd_parse(int key, char *arg, struct argp_state *st)
{
  struct cpd_args *a = (struct cpd_args *)st->input;
  if (arg && arg[0] == '=') ++arg;
  switch (key) {
  case 'i': a->niters = (uint32_t)atoi(arg); break;
  case OPT_TOL: a->tol = atof(arg); break;
  case 'r': a->rank = (uint32_t)atoi(arg); break;
  case 't': a->nthreads = (uint32_t)atoi(arg); break;
  case OPT_NOWRITE: a->write = 0; break;
  case 'v': a->verbose += 1; break;
  case OPT_TILE: break;
  case OPT_PREC: a->prec = parse_prec(arg); break;
  case OPT_SEED: a->seed = strtoull(arg, NULL, 10); break;
  case OPT_PLANNER: a->planner = parse_planner(arg); break;
  case ARGP_KEY_ARG:
    if (a->tensor) argp_usage(st);
    a->tensor = arg;
    break;
  case ARGP_KEY_END:
    if (!a->tensor) argp_usage(st);
    break;
  default: return ARGP_ERR_UNKNOWN;
  }
  return 0;
}

static void write_matrix(char const *fname, double const *A, uint64_t I,
                         uint64_t J)
{
  FILE *f = fopen(fname, "w");
  if (!f) return;
  for (uint64_t i = 0; i < I; ++i) {
    for (uint64_t j = 0; j < J; ++j) fprintf(f, "%+0.8e ", A[i * J + j]);
    fprintf(f, "\n");
  }
  fclose(f);
}

static int cmd_cpd(int argc, char **argv)
{
  struct cpd_args a = {NULL, 50, 10, 0, 1e-5, 1, 0, SPLATT_ADATM_FP64,
                       (uint64_t)time(NULL), SPLATT_ADATM_PLAN_ANALYTIC};
  struct argp ap = {cpd_opts, cpd_parse, "TENSOR",
                    "splatt-cpd -- Compute the CPD of a sparse tensor.", 0, 0, 0};
  argp_parse(&ap, argc, argv, ARGP_IN_ORDER, 0, &a);
  printf("Use splatt_cpd_cmd_adaptive.\n");
  adatm_coo t;
  if (load_tensor(a.tensor, &t) != 0) {
    fprintf(stderr, "SPLATT: cannot load '%s': %s\n", a.tensor, adatm_last_error());
    return EXIT_FAILURE;
  }
  print_header(&t, a.tensor);
  adatm_plan plan;
  if (plan_for(&t, a.rank, a.planner, a.prec, &plan) != 0) return EXIT_FAILURE;
  adatm_engine *e = NULL;
  if (adatm_engine_create(&t, a.rank, &plan, a.prec, 0, &e) != SPLATT_SUCCESS) {
    fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
    return EXIT_FAILURE;
  }
  print_storage(e, &plan, t.nmodes);
  uint32_t ug[ADATM_MAX_NMODES], ut[ADATM_MAX_NMODES];
  adatm_plan_use(&plan, t.nmodes, ug, ut);
  printf("use_csfs: ");
  for (uint32_t m = 0; m < t.nmodes; ++m) printf("%u ", ug[m]);
  printf("\nuse_tags: ");
  for (uint32_t m = 0; m < t.nmodes; ++m) printf("%u ", ut[m]);
  printf("\n");
  adatm_engine_init_factors(e, a.seed);
  if (adatm_engine_cpd_begin(e) != SPLATT_SUCCESS) return EXIT_FAILURE;
  double fit = 0.0, oldfit = 0.0;
  double const t_all = now();
  for (uint32_t it = 0; it < a.niters; ++it) {
    double const t0 = now();
    if (adatm_engine_cpd_sweep(e, it, 1, &fit) != SPLATT_SUCCESS) {
      fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
      return EXIT_FAILURE;
    }
    printf("  its = %3u (%0.3fs)  fit = %0.5f  delta = %+0.4e\n", it + 1,
           now() - t0, fit, fit - oldfit);
    if (it > 0 && fabs(fit - oldfit) < a.tol) break;
    oldfit = fit;
  }
  printf("CPD = %0.3fs\n", now() - t_all);
  printf("Final fit: %f\n", fit);
  if (a.write) {
    double *lam = (double *)malloc(a.rank * sizeof(double));
    adatm_engine_get_lambda(e, lam);
    /* unit 2-norm columns with the norms folded into lambda */
    double *A[ADATM_MAX_NMODES];
    for (uint32_t m = 0; m < t.nmodes; ++m) {
      A[m] = (double *)malloc(t.dims[m] * a.rank * sizeof(double));
      adatm_engine_get_factor(e, m, A[m]);
      for (uint32_t r = 0; r < a.rank; ++r) {
        double s = 0.0;
        for (uint64_t i = 0; i < t.dims[m]; ++i) s += A[m][i * a.rank + r] * A[m][i * a.rank + r];
        s = sqrt(s);
        if (s > 0.0)
          for (uint64_t i = 0; i < t.dims[m]; ++i) A[m][i * a.rank + r] /= s;
        lam[r] *= s;
      }
    }
    FILE *f = fopen("lambda.mat", "w");
    if (f) {
      for (uint32_t r = 0; r < a.rank; ++r) fprintf(f, "%0.8e\n", lam[r]);
      fclose(f);
    }
    for (uint32_t m = 0; m < t.nmodes; ++m) {
      char fn[64];
      snprintf(fn, sizeof(fn), "mode%u.mat", m + 1);
      write_matrix(fn, A[m], t.dims[m], a.rank);
      free(A[m]);
    }
    free(lam);
  }
  adatm_engine_destroy(e);
  adatm_coo_free(&t);
  return EXIT_SUCCESS;
}

int main(int argc, char **argv)
{
  if (argc < 2) {
    fprintf(stderr, "usage: splatt <bench|cpd> TENSOR [options]\n");
    return EXIT_FAILURE;
  }
  if (!strcmp(argv[1], "bench")) return cmd_bench(argc - 1, argv + 1);
  if (!strcmp(argv[1], "cpd")) return cmd_cpd(argc - 1, argv + 1);
  fprintf(stderr, "splatt: unknown command '%s' (bench, cpd)\n", argv[1]);
  return EXIT_FAILURE;
}
