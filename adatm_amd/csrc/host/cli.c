/*
 * `splatt` command-line tool: the AdaTM subcommands on one MI355X.
 *
 *   splatt bench TENSOR -a adatm [-i NITERS] [-r RANK] [-t NTHREADS] [--precision P] [--planner PL]
 *   splatt cpd   TENSOR [-i NITERS] [--tol TOL] [-r RANK] [-t NTHREADS]
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
  int adatm;
  char *alg;
  uint32_t niters, nthreads, rank;
  int prec, planner;
};

static struct argp_option bench_opts[] = {
    {"alg", 'a', "ALG", 0, "algorithm to benchmark (adatm)", 0},
    {"iters", 'i', "NITERS", 0, "number of iterations to use (default: 5)", 0},
    {"threads", 't', "NTHREADS", 0, "host threads (default: 1)", 0},
    {"rank", 'r', "RANK", 0, "rank of decomposition to find (default: 16)", 0},
    {"precision", 'p', "P", 0, "fp64 | fp32 | bf16rows (default: fp64)", 0},
    {"planner", 'P', "PLANNER", 0, "analytic | learned | exhaustive (default: analytic)", 0},
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
    a->alg = arg;
    a->adatm = !strcmp(arg, "adatm");
    break;
  case 'i': a->niters = (uint32_t)atoi(arg); break;
  case 't': a->nthreads = (uint32_t)atoi(arg); break;
  case 'r': a->rank = (uint32_t)atoi(arg); break;
  case 'p': a->prec = parse_prec(arg); break;
  case 'P': a->planner = parse_planner(arg); break;
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

static int cmd_bench(int argc, char **argv)
{
  struct bench_args a = {NULL, 0, NULL, 5, 1, 16, SPLATT_ADATM_FP64,
                         SPLATT_ADATM_PLAN_ANALYTIC};
  struct argp ap = {bench_opts, bench_parse, "TENSOR -a adatm",
                    "splatt-bench -- benchmark the memoized MTTKRP sequence.", 0, 0, 0};
  argp_parse(&ap, argc, argv, ARGP_IN_ORDER, 0, &a);
  if (!a.adatm) {
    fprintf(stderr, "SPLATT: algorithm '%s' is not recognized (only 'adatm').\n",
            a.alg ? a.alg : "");
    return EXIT_FAILURE;
  }
  adatm_coo t;
  if (load_tensor(a.tensor, &t) != 0) {
    fprintf(stderr, "SPLATT: cannot load '%s': %s\n", a.tensor, adatm_last_error());
    return EXIT_FAILURE;
  }
  print_header(&t, a.tensor);
  printf("Benchmarking ---------------------------------------------------\n");
  printf("RANK=%u ITS=%u\n", a.rank, a.niters);
  adatm_plan plan;
  if (plan_for(&t, a.rank, a.planner, a.prec, &plan) != 0) return EXIT_FAILURE;
  adatm_engine *e = NULL;
  if (adatm_engine_create(&t, a.rank, &plan, a.prec, 0, &e) != SPLATT_SUCCESS) {
    fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
    return EXIT_FAILURE;
  }
  print_storage(e, &plan, t.nmodes);
  adatm_engine_init_factors(e, 1);
  for (uint32_t i = 0; i < a.niters; ++i) {
    double const t0 = now();
    for (uint32_t m = 0; m < t.nmodes; ++m) {
      double const tm = now();
      if (adatm_engine_mttkrp(e, m, NULL) != SPLATT_SUCCESS) {
        fprintf(stderr, "SPLATT: %s\n", adatm_last_error());
        return EXIT_FAILURE;
      }
      printf("  mode %u %0.3fs\n", m + 1, now() - tm);
    }
    printf("    its = %3u (%0.3fs)\n", i + 1, now() - t0);
  }
  adatm_engine_destroy(e);
  adatm_coo_free(&t);
  return EXIT_SUCCESS;
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
    {"threads", 't', "NTHREADS", 0, "host threads (default: #cores)", 0},
    {"nowrite", OPT_NOWRITE, 0, 0, "do not write output to file", 0},
    {"verbose", 'v', 0, 0, "turn on verbose output (default: no)", 0},
    {"precision", 'p', "P", 0, "fp64 | fp32 | bf16rows (default: fp64)", 0},
    {"seed", OPT_SEED, "SEED", 0, "factor initialisation seed (default: time)", 0},
    {"planner", 'P', "PLANNER", 0, "analytic | learned | exhaustive (default: analytic)", 0},
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
  case 'p': a->prec = parse_prec(arg); break;
  case OPT_SEED: a->seed = strtoull(arg, NULL, 10); break;
  case 'P': a->planner = parse_planner(arg); break;
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
