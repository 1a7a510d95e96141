/*
 * splatt.h -- public C API of the MI355X AdaTM library (libsplatt_adatm.so).
 *
 * Drop-in for SPLATT v1.1.1 + the AdaTM patch (hpcgarage/AdaTM).  Every type,
 * enum value and struct layout below matches include/splatt.h of the
 * reference so that a C program written against it compiles and links
 * unchanged.  The memoized MTTKRP / CP-ALS path behind these entry points
 * runs on one MI355X (gfx950) through hand-written HIP kernels.
 *
 * Additions (all appended, no existing value moves):
 *   - options SPLATT_OPTION_ADATM_* after SPLATT_OPTION_COMM
 *     (SPLATT_OPTION_NOPTIONS grows; arrays come from splatt_default_opts()),
 *   - error codes SPLATT_ERROR_DEVICE, SPLATT_ERROR_NUMERIC after
 *     SPLATT_ERROR_NOMEMORY,
 *   - sptensor_t / matrix_t, which SPLATT keeps in internal headers,
 *   - declarations of the AdaTM entry points that the reference defines but
 *     never declares publicly (splatt_cpd_als_adaptive, planner functions).
 */
#ifndef SPLATT_SPLATT_H
#define SPLATT_SPLATT_H

#include <inttypes.h>
#include <float.h>
#include <stddef.h>
#include <stdio.h>

/* ----------------------------------------------------------------------- *
 * Types (reference include/splatt.h:26-68; 64-bit defaults)
 * ----------------------------------------------------------------------- */
#ifndef SPLATT_IDX_TYPEWIDTH
#define SPLATT_IDX_TYPEWIDTH 64
#endif
#ifndef SPLATT_VAL_TYPEWIDTH
#define SPLATT_VAL_TYPEWIDTH 64
#endif
#if SPLATT_IDX_TYPEWIDTH != 64 || SPLATT_VAL_TYPEWIDTH != 64
#error "libsplatt_adatm is built with 64-bit splatt_idx_t and splatt_val_t"
#endif
typedef uint64_t splatt_idx_t;
typedef double splatt_val_t;
#define SPLATT_IDX_MAX UINT64_MAX
#define SPLATT_PF_IDX PRIu64
#define SPLATT_VAL_MIN DBL_MIN
#define SPLATT_VAL_MAX DBL_MAX
#define SPLATT_PF_VAL "f"

#define SPLATT_VER_MAJOR 1
#define SPLATT_VER_MINOR 1
#define SPLATT_VER_SUBMINOR 1

/* reference include/splatt.h:84-87 */
#ifndef SPLATT_MAX_NMODES
#define SPLATT_MAX_NMODES ((splatt_idx_t)88)
#endif

/* reference include/splatt.h:93-111, extended at the end */
typedef enum {
  SPLATT_OPTION_NTHREADS,
  SPLATT_OPTION_TOLERANCE,
  SPLATT_OPTION_NITER,
  SPLATT_OPTION_VERBOSITY,
  SPLATT_OPTION_RANDSEED,
  SPLATT_OPTION_CSF_ALLOC,
  SPLATT_OPTION_TILE,
  SPLATT_OPTION_TILEDEPTH,
  SPLATT_OPTION_DECOMP,
  SPLATT_OPTION_COMM,
  /* --- MI355X AdaTM additions --- */
  SPLATT_OPTION_ADATM_PRECISION, /* splatt_adatm_precision */
  SPLATT_OPTION_ADATM_PLANNER,   /* splatt_adatm_planner */
  SPLATT_OPTION_ADATM_GRAPH,     /* 1: capture each ALS sweep in one hipGraph */
  SPLATT_OPTION_ADATM_REFQUIRKS, /* 1: analytic model keeps the reference's
                                    statistics reuse (adatm_bench.c:246-262) */
  SPLATT_OPTION_ADATM_SPACE_LIMIT, /* bytes; reference SLIMIT = 90e9 */
  SPLATT_OPTION_ADATM_DEVICE,    /* GPU ordinal */
  SPLATT_OPTION_ADATM_RANK_HINT, /* rank the learned / exhaustive planners
                                    plan for (CSFs are built before the rank
                                    of a decomposition is known) */
  SPLATT_OPTION_NOPTIONS
} splatt_option_type;

/* reference include/splatt.h:117-125, extended */
typedef enum {
  SPLATT_SUCCESS = 1,
  SPLATT_ERROR_BADINPUT,
  SPLATT_ERROR_NOMEMORY,
  SPLATT_ERROR_DEVICE, /* HIP runtime error on the GPU path */
  SPLATT_ERROR_NUMERIC /* CP-ALS normal equations not solvable (no positive pivot) */
} splatt_error_type;

typedef enum {
  SPLATT_VERBOSITY_NONE,
  SPLATT_VERBOSITY_LOW,
  SPLATT_VERBOSITY_HIGH,
  SPLATT_VERBOSITY_MAX
} splatt_verbosity_type;

typedef enum {
  SPLATT_NOTILE,
  SPLATT_DENSETILE,
  SPLATT_SYNCTILE,
  SPLATT_COOPTILE,
} splatt_tile_type;

/* reference include/splatt.h:156-164 */
typedef enum {
  SPLATT_CSF_ONEMODE,
  SPLATT_CSF_TWOMODE,
  SPLATT_CSF_ALLMODE,
  SPLATT_ADAPTIVE,
  SPLATT_ADAPTIVE_TIME_EFFICIENT,
  SPLATT_ADAPTIVE_SPACE_EFFICIENT,
} splatt_csf_type;

typedef enum {
  SPLATT_DECOMP_COARSE,
  SPLATT_DECOMP_MEDIUM,
  SPLATT_DECOMP_FINE
} splatt_decomp_type;

typedef enum {
  SPLATT_COMM_POINT2POINT,
  SPLATT_COMM_ALL2ALL
} splatt_comm_type;

/* Arithmetic of the GPU path. */
typedef enum {
  SPLATT_ADATM_FP64 = 0,     /* factors, memo, accumulation in FP64 */
  SPLATT_ADATM_FP32 = 1,     /* FP32 factors/memo/accumulation; FP64 solve+fit */
  SPLATT_ADATM_BF16ROWS = 2  /* FP32 master factors, BF16 copies gathered by
                                the MTTKRP kernels, FP32 accumulation */
} splatt_adatm_precision;

typedef enum {
  SPLATT_ADATM_PLAN_ANALYTIC = 0,   /* the reference's model (min pops) */
  SPLATT_ADATM_PLAN_LEARNED = 1,    /* MLP cost model, analytic fallback */
  SPLATT_ADATM_PLAN_EXHAUSTIVE = 2  /* time every candidate on the GPU */
} splatt_adatm_planner;

static double const SPLATT_VAL_OFF = -DBL_MAX;

/* ----------------------------------------------------------------------- *
 * Data structures (reference include/splatt.h:209-358, same layouts)
 * ----------------------------------------------------------------------- */
typedef struct splatt_kruskal {
  splatt_idx_t rank;
  splatt_val_t *factors[SPLATT_MAX_NMODES];
  splatt_val_t *lambda;
  splatt_idx_t nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  double fit;
} splatt_kruskal;

typedef struct {
  splatt_idx_t nfibs[SPLATT_MAX_NMODES];
  splatt_idx_t *fptr[SPLATT_MAX_NMODES];
  splatt_idx_t *fids[SPLATT_MAX_NMODES];
  splatt_val_t *vals;
} csf_sparsity;

typedef struct splatt_csf {
  splatt_idx_t nnz;
  splatt_idx_t nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  splatt_idx_t dim_perm[SPLATT_MAX_NMODES];
  splatt_tile_type which_tile;
  splatt_idx_t ntiles;
  splatt_idx_t tile_dims[SPLATT_MAX_NMODES];
  csf_sparsity *pt;
} splatt_csf;

typedef struct rcsf_seq_adaptive {
  splatt_idx_t begin_imten;
  splatt_idx_t n_imten;
  splatt_csf *rcsfs;
} rcsf_seq_adaptive;

typedef struct group_properties {
  splatt_idx_t memo_mode;
  splatt_idx_t begin_imten;
  splatt_idx_t n_imten;
  splatt_idx_t *o_opt;
} group_properties;

typedef struct configurations_adaptive {
  splatt_idx_t n_grp;
  group_properties *grp_prop;
  size_t pspace;
  size_t pops;
  double ptime;
  double rtime;
} configurations_adaptive;

typedef struct pair {
  splatt_idx_t x;
  splatt_idx_t y;
} pair;

/*
 * Coordinate tensor.  SPLATT keeps sptensor_t in its internal sptensor.h,
 * which is not distributed with the reference; these are the fields the
 * AdaTM functions read (tt->nmodes, tt->nnz, tt->dims, tt->ind, tt->vals).
 */
typedef struct {
  splatt_idx_t nnz;
  splatt_idx_t nmodes;
  splatt_idx_t *dims;
  splatt_idx_t **ind;
  splatt_val_t *vals;
  int tiled;
  splatt_idx_t *indmap[SPLATT_MAX_NMODES]; /* compacted id -> original id */
} sptensor_t;

/* Dense row-major matrix (SPLATT matrix.h field names). */
typedef struct {
  splatt_idx_t I;
  splatt_idx_t J;
  splatt_val_t *vals;
  int rowmajor;
} matrix_t;

/*
 * Index relabelling of a tensor (SPLATT reorder.h, not distributed with the
 * reference; used at cmd_bench.c:119-149, 243-251 and adatm_bench.c:26-51,
 * 529, 654).  perms[m][old] = new, iperms[m][new] = old; a NULL perms[m]
 * leaves mode m as it is.
 */
typedef enum {
  PERM_GRAPH,    /* partition of the m-partite index graph (all modes) */
  PERM_HGRAPH,   /* partition of the fibers of one mode's CSF */
  PERM_FIBSCHED, /* partition of the slices of one mode's CSF */
  PERM_ERROR
} splatt_perm_type;

typedef struct {
  splatt_idx_t *perms[SPLATT_MAX_NMODES];
  splatt_idx_t *iperms[SPLATT_MAX_NMODES];
} permutation_t;

#define SPLATT_ADATM_SLIMIT 90000000000UL /* reference adatm_base.h:4 */

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------- *
 * Options (reference include/splatt.h:384-391)
 * ----------------------------------------------------------------------- */
double *splatt_default_opts(void);
void splatt_free_opts(double *opts);

/* ----------------------------------------------------------------------- *
 * Data structures (reference include/splatt.h:416-467)
 *
 * With opts[SPLATT_OPTION_CSF_ALLOC] = SPLATT_ADAPTIVE or
 * SPLATT_ADAPTIVE_TIME_EFFICIENT the returned array holds one CSF per
 * memoization group, and the plan, the vCSF memo buffers and the device copy
 * live in a side handle keyed by the returned pointer;
 * ADAPTIVE_SPACE_EFFICIENT keeps the CSFs of the memoising groups only and
 * computes the other groups' modes from their leaves.  ONEMODE/TWOMODE/ALLMODE
 * return the plain SPLATT layouts (1, 2 or nmodes CSFs).  Coordinates are
 * taken as given (duplicates stay separate nonzeros; dims = max index + 1).
 * ----------------------------------------------------------------------- */
int splatt_csf_load(char const *const fname, splatt_idx_t *nmodes,
                    splatt_csf **tensors, double const *const options);

int splatt_csf_convert(splatt_idx_t const nmodes, splatt_idx_t const nnz,
                       splatt_idx_t **const inds, splatt_val_t *const vals,
                       splatt_csf **tensors, double const *const options);

void splatt_free_csf(splatt_csf *tensors, double const *const options);

void splatt_free_kruskal(splatt_kruskal *factored);

/* SPLATT-internal allocator named by the AdaTM CLI (cmd_cpd.c:292). */
splatt_csf *splatt_csf_alloc(sptensor_t *const tt, double const *const opts);

/* ----------------------------------------------------------------------- *
 * Factorization / operations (reference include/splatt.h:493-525)
 * ----------------------------------------------------------------------- */
int splatt_cpd_als(splatt_csf const *const tensors,
                   splatt_idx_t const nfactors, double const *const options,
                   splatt_kruskal *factored);

int splatt_mttkrp(splatt_idx_t const mode, splatt_idx_t const ncolumns,
                  splatt_csf const *const tensors, splatt_val_t **matrices,
                  splatt_val_t *const matout, double const *const options);

/* ----------------------------------------------------------------------- *
 * AdaTM entry points (reference adatm_bench.h:22-110, adatm_rcsf.h:18-161,
 * mttkrp.h:42-62, adatm_cpd.c:205-215).
 * ----------------------------------------------------------------------- */
void decide_product_order(sptensor_t *const tt, splatt_idx_t **product_order);

splatt_idx_t decide_candidate_configs(sptensor_t *const tt,
                                      splatt_idx_t const **product_order,
                                      configurations_adaptive **configs);

void predict_candidates_adaptive(sptensor_t *const tt,
                                 splatt_idx_t const nfactors,
                                 configurations_adaptive *configs,
                                 splatt_idx_t const nconfigs);

splatt_idx_t limit_candidates_by_space(
    configurations_adaptive const *const configs, splatt_idx_t const nconfigs,
    splatt_idx_t **const real_config_locs);

splatt_idx_t choose_candidates_by_strategy(
    configurations_adaptive const *const configs, splatt_idx_t const nconfigs,
    splatt_idx_t const *const real_config_locs,
    splatt_idx_t const n_real_configs, int strategy);

splatt_idx_t decide_parameters_auto(sptensor_t *const tt,
                                    splatt_idx_t const nfactors, int strategy,
                                    splatt_idx_t const **product_order,
                                    group_properties **grp_prop);

void decide_use_csfs(splatt_idx_t const nmodes,
                     group_properties *const grp_prop, int const n_grp,
                     splatt_idx_t const n_csf, splatt_idx_t *use_csfs,
                     splatt_idx_t *use_tags);

splatt_csf *csf_alloc_adaptive(sptensor_t *const tt,
                               group_properties *const grp_prop,
                               splatt_idx_t const n_grp,
                               double const *const opts, splatt_idx_t *n_csf);

rcsf_seq_adaptive *rcsf_alloc_adaptive(splatt_csf *const ct,
                                       splatt_idx_t const nfactors,
                                       group_properties *const grp_prop,
                                       splatt_idx_t const n_grp,
                                       splatt_idx_t const n_csf,
                                       double const *const opts,
                                       splatt_idx_t *n_rcsf);

size_t csf_storage_adaptive(splatt_csf const *const tensors,
                            splatt_idx_t const n_csf,
                            group_properties *const grp_prop, int const n_grp);

size_t rcsf_storage_adaptive(rcsf_seq_adaptive const *const seq_rcsf,
                             splatt_idx_t const n_rcsf);

void count_nfibs(sptensor_t const *const tt, splatt_idx_t const mode,
                 splatt_idx_t const *const o_opt, splatt_idx_t *const fprev,
                 splatt_idx_t *const nfibs_per_grp, splatt_idx_t **const fnext);

size_t predict_csf_bytes_adaptive(sptensor_t *const tt,
                                  splatt_idx_t const *const nfibs_per_grp);
size_t predict_rcsf_bytes_adaptive(sptensor_t *const tt,
                                   splatt_idx_t const nfactors,
                                   splatt_idx_t const begin_imten,
                                   splatt_idx_t const n_imten,
                                   splatt_idx_t const *const nfibs_per_grp);
size_t predict_csf_ops_adaptive(sptensor_t *const tt,
                                splatt_idx_t const nfactors,
                                splatt_idx_t const *const nfibs_per_grp);
size_t predict_rcsf_ops_adaptive(sptensor_t *const tt,
                                 splatt_idx_t const nfactors,
                                 splatt_idx_t const begin_imten,
                                 splatt_idx_t const n_imten,
                                 splatt_idx_t const num_reMTTKRP,
                                 splatt_idx_t const *const nfibs_per_grp);

void rcsf_reverse_mode_order(splatt_idx_t const begin, splatt_idx_t const mode,
                             splatt_idx_t *const rdims);

void pair_sort(pair *const pair_list, splatt_idx_t const size);
splatt_idx_t argmin_elem_range(splatt_idx_t const *const arr,
                               splatt_idx_t const N, splatt_idx_t const con);
splatt_idx_t array_min_range(splatt_idx_t const *const arr,
                             splatt_idx_t const N, splatt_idx_t const con);

/*
 * One memoized MTTKRP of the sequence (reference mttkrp.c:2037-2090).  mats
 * has SPLATT_MAX_NMODES+1 entries; the output is mats[SPLATT_MAX_NMODES].
 * `thds` is accepted for signature compatibility and ignored (the GPU needs
 * no per-thread scratch).  `tensors` must come from csf_alloc_adaptive or
 * splatt_csf_convert with an adaptive CSF_ALLOC.
 */
void mttkrp_csf_adaptive(splatt_csf const *const tensors,
                         rcsf_seq_adaptive const *const rs_seq,
                         splatt_idx_t const n_csf, matrix_t **mats,
                         splatt_idx_t const mode, void *const thds,
                         group_properties *const grp_prop, int const n_grp,
                         splatt_idx_t const use_csf, splatt_idx_t const use_tag,
                         double const *const opts);

/* SPLATT's MPI rank information (splatt_mpi.h, not distributed with the
 * reference); accepted for signature compatibility and ignored. */
typedef struct rank_info rank_info;

/* The ALS loop (reference adatm_cpd.h:29-41, adatm_cpd.c:256-374) on caller
 * matrices: mats[m] (dims[m] x nfactors) hold the initial factors and
 * receive the final ones (unit 2-norm columns, the norms in lambda); the
 * plan must be the one the CSFs were built with.  Returns the final fit (NaN
 * on failure, message in adatm_last_error()). */
double cpd_als_iterate_adaptive(splatt_csf const *const tensors,
                                rcsf_seq_adaptive *const rs_seq,
                                splatt_idx_t const n_csf,
                                group_properties const *const grp_prop,
                                splatt_idx_t const n_grp,
                                splatt_idx_t const *const use_csfs,
                                splatt_idx_t const *const use_tags, matrix_t **mats,
                                splatt_val_t *const lambda, splatt_idx_t const nfactors,
                                rank_info *const rinfo, double const *const opts);

/* SPLATT-internal helpers the AdaTM drivers call (cmd_bench.c:234-266,
 * cmd_cpd.c:139-154, 255, adatm_bench.c:650): tt_read reads the coordinate
 * text format of splatt_csf_load (dims = largest index + 1); csf_free and
 * splatt_csf_free are splatt_free_csf; mat_alloc gives a zeroed row-major
 * matrix. */
sptensor_t *tt_read(char const *const fname);
void tt_free(sptensor_t *tt);
void csf_free(splatt_csf *tensors, double const *const opts);
void splatt_csf_free(splatt_csf *tensors, double const *const opts);
matrix_t *mat_alloc(splatt_idx_t const nrows, splatt_idx_t const ncols);
void mat_free(matrix_t *M);

/* Relabelling (SPLATT reorder.h; callers cmd_bench.c:243-251,
 * adatm_bench.c:26-51).
 *   tt_perm     reads the partition file PFILE (one part id per vertex, in
 *               vertex order, whitespace separated), builds the relabelling
 *               of `type` (basis mode `mode` for PERM_HGRAPH/PERM_FIBSCHED)
 *               and applies it to tt.  NULL on error (adatm_last_error()).
 *   perm_alloc  identity permutations for the first nmodes modes (nmodes 0:
 *               every mode NULL, "no relabelling", cmd_bench.c:250).
 *   perm_apply  tt->ind[m][j] = perm[m][tt->ind[m][j]] for every non-NULL
 *               perm[m].
 *   perm_matrix row perm[i] of the result is row i of mat (the rows follow
 *               their index to its new label); retmat NULL allocates. */
permutation_t *tt_perm(sptensor_t *const tt, splatt_perm_type const type,
                       splatt_idx_t const mode, char const *const pfile);
permutation_t *perm_alloc(splatt_idx_t const *const dims, splatt_idx_t const nmodes);
void perm_free(permutation_t *perm);
void perm_apply(sptensor_t *const tt, splatt_idx_t **perm);
matrix_t *perm_matrix(matrix_t const *const mat, splatt_idx_t const *const perm,
                      matrix_t *retmat);

/* CP-ALS over a memoization plan (reference adatm_cpd.c:205-253). */
int splatt_cpd_als_adaptive(splatt_csf const *const tensors,
                            rcsf_seq_adaptive *const rs_seq,
                            splatt_idx_t const n_csf,
                            group_properties const *const grp_prop,
                            splatt_idx_t const n_grp,
                            splatt_idx_t const *const use_csfs,
                            splatt_idx_t const *const use_tags,
                            splatt_idx_t const nfactors,
                            double const *const options,
                            splatt_kruskal *factored);

#ifdef __cplusplus
}
#endif

#endif
