/*
 * adatm.h -- native entry points of libsplatt_adatm.so beyond the SPLATT API.
 *
 * These expose the device-resident engine directly (no host CSF copies), the
 * synthetic workload generators and the planner's statistics.  All functions
 * return a splatt_error_type (SPLATT_SUCCESS == 1) unless stated otherwise.
 * Plain C ABI: pointers, sizes, no C++ or torch types.
 */
#ifndef ADATM_ADATM_H
#define ADATM_ADATM_H

#include <stdint.h>
#include <stddef.h>
#include "splatt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ADATM_MAX_NMODES 16 /* engine limit; the planner handles N = 3..16 */

/* ----------------------------------------------------------------------- *
 * Coordinate tensors in host memory.
 * ----------------------------------------------------------------------- */
typedef struct {
  uint32_t nmodes;
  uint64_t nnz;
  uint64_t dims[ADATM_MAX_NMODES];
  uint32_t *ind[ADATM_MAX_NMODES]; /* nnz entries each, 0-based */
  double *vals;                    /* nnz entries */
} adatm_coo;

/* Index distributions of the synthetic generators. */
typedef enum {
  ADATM_DIST_UNIFORM = 0,
  ADATM_DIST_ZIPF = 1,     /* P(k) ~ (k+1)^-s on [0, dim) by exact inverse CDF */
  ADATM_DIST_CLUSTER = 2   /* Gaussian clusters, see adatm_synth_spec */
} adatm_index_dist;

typedef enum {
  ADATM_VALS_UNIFORM01 = 0,  /* U(0,1) */
  ADATM_VALS_LOGNORMAL = 1,  /* exp(N(0,1)) */
  ADATM_VALS_UNIFORM_PM1 = 2 /* U(-1,1) */
} adatm_value_dist;

typedef struct {
  uint32_t nmodes;
  uint64_t dims[ADATM_MAX_NMODES];
  uint64_t ndraws;         /* coordinates drawn (before duplicate merging) */
  int index_dist;          /* adatm_index_dist, all modes */
  double zipf_s[ADATM_MAX_NMODES];
  uint64_t nclusters;      /* ADATM_DIST_CLUSTER */
  double cluster_sigma;    /* ADATM_DIST_CLUSTER */
  int value_dist;          /* adatm_value_dist */
  uint64_t seed;
} adatm_synth_spec;

/* Draw a raw coordinate tensor (duplicates possible) from a spec.
 * Philox4x32-10 counter stream: draw j only depends on (seed, j). */
int adatm_synth(adatm_synth_spec const *spec, adatm_coo *out);

/* Fill the spec of one of the named workloads:
 *   1: 4-mode 2000x1500x1000x500, 2M uniform draws, U(0,1) values
 *   2: 5-mode 1e5x5e4x2e4x5e3x1e3, 100M draws, Zipf(1.1) per mode, lognormal
 *   3: 8-mode 1000^8, 50M draws, 100000 Gaussian clusters (sigma 20), U(-1,1)
 * ndraws_override > 0 replaces the number of draws (bounded samples). */
int adatm_synth_named(int config, uint64_t ndraws_override,
                      adatm_synth_spec *spec);

/* Merge duplicate coordinates (values summed in draw order), drop empty
 * indices of every mode (dims shrink; indmap[m][new] = old, optional) and
 * leave the tensor sorted lexicographically by mode 0,1,...,N-1. */
int adatm_coo_canonicalize(adatm_coo *t, uint32_t **indmap_out);

void adatm_coo_free(adatm_coo *t);

/* malloc an nnz-entry tensor with the given dims (contents undefined) */
int adatm_coo_alloc(uint32_t nmodes, uint64_t nnz, uint64_t const *dims,
                    adatm_coo *out);

/* Relabelling of a coordinate tensor's indices (the computation behind
 * SPLATT's tt_perm).  type is splatt_perm_type (PERM_GRAPH, PERM_HGRAPH,
 * PERM_FIBSCHED); `mode` is the basis mode of the latter two, whose CSF
 * (root `mode`, the other modes shortest first) is built on the GPU
 * `device` (device < 0: on the host).  parts holds one part id per vertex:
 *   PERM_GRAPH     sum(dims) vertices, mode 0's indices first;
 *   PERM_HGRAPH    the CSF's fibers (nodes of level N-2), in CSF order;
 *   PERM_FIBSCHED  the CSF's slices (nodes of level 0), in CSF order.
 * PERM_GRAPH orders each mode's indices by (part, index).  The other two
 * visit the vertices by (part, CSF position) and give every mode's indices
 * new labels in the order the visit first touches them (the vertex's path
 * from the root, then its subtree in CSF order); untouched indices follow
 * in index order.  nparts holds at least that many part ids (the first
 * ones are used; fewer is SPLATT_ERROR_BADINPUT).  perms[m] (dims[m]
 * entries each, caller-allocated) gets old -> new. */
int adatm_perm_compute(adatm_coo const *t, int type, uint32_t mode,
                       uint32_t const *parts, uint64_t nparts, int device,
                       uint32_t **perms);
/* number of vertices adatm_perm_compute reads for (type, mode): sum(dims),
 * or the fiber / slice count of the basis CSF (counted on `device`) */
int adatm_perm_nvtxs(adatm_coo const *t, int type, uint32_t mode, int device,
                     uint64_t *nvtxs);
/* read a partition file: whitespace-separated non-negative integers below
 * 2^32 (anything else is SPLATT_ERROR_BADINPUT); malloc'd result */
int adatm_part_read(char const *fname, uint64_t *count, uint32_t **parts);
/* t->ind[m][j] = perms[m][t->ind[m][j]] for every non-NULL perms[m] */
void adatm_coo_relabel(adatm_coo *t, uint32_t *const *perms);

/* Factor matrix initialisation used by both the CPU and the GPU paths:
 * A[i*R + r] = U(0,1) from Philox stream (seed, mode). */
void adatm_rand_matrix(uint64_t seed, uint32_t mode, uint64_t rows,
                       uint64_t rank, double *out);

/* Philox4x32-10 block (exposed for known-answer tests). */
void adatm_philox4x32_10(uint32_t const ctr[4], uint32_t const key[2],
                         uint32_t out[4]);

/* ----------------------------------------------------------------------- *
 * Memoization plans.
 * ----------------------------------------------------------------------- */
typedef struct {
  uint32_t memo_mode;
  uint32_t begin_imten;
  uint32_t n_imten;
  uint32_t o_opt[ADATM_MAX_NMODES];
} adatm_group;

typedef struct {
  uint32_t n_grp;
  adatm_group grp[ADATM_MAX_NMODES];
  uint64_t pspace; /* analytic model: predicted bytes (reference units) */
  uint64_t pops;   /* analytic model: predicted products */
} adatm_plan;

/* Enumerate the candidate plans of an N-mode tensor (adatm_bench.c:276-458);
 * returns the count (<= maxc written).  N = 3 yields the one valid plan. */
int adatm_plan_candidates(uint32_t nmodes, uint64_t const *dims,
                          adatm_plan *cands, int maxc);

/* Number of distinct length-(l+1) prefixes of the coordinates under the
 * level order `order`, for every level l (nfibs[N-1] = nnz).  Runs on the
 * GPU (radix sort + boundary scan) when device >= 0, else on the host. */
int adatm_count_fibers(adatm_coo const *t, uint32_t const *order,
                       uint64_t *nfibs, int device);

/* The analytic planner (decide_parameters_auto, strategy 1 = min pops).
 * Fills every candidate's pspace/pops, returns the chosen index in *chosen.
 * ref_quirks = 1 keeps the reference's statistics reuse between candidates
 * with equal group counts (adatm_bench.c:246-262). */
int adatm_plan_analytic(adatm_coo const *t, uint64_t rank, int ref_quirks,
                        double space_limit, int device, adatm_plan *cands,
                        int maxc, int *ncands, int *chosen);

/* SPACE_EFFICIENT form of a plan (adatm_rcsf.c:528-552): the memoising
 * groups only, and the use maps that serve every other group's modes from the
 * leaves of the nearest memoising group before it (tag 2; mttkrp.c:1976-2029
 * with D7/D8 corrected).  Returns the number of CSFs kept. */
int adatm_plan_space_efficient(adatm_plan const *p, uint32_t nmodes,
                               adatm_plan *out, uint32_t *use_grp,
                               uint32_t *use_tag);

/* Map each mode to (group, tag): tag 0 = recomputed root, 1 = memo reuse
 * (mttkrp.c:1958-1974). */
void adatm_plan_use(adatm_plan const *p, uint32_t nmodes, uint32_t *use_grp,
                    uint32_t *use_tag);

/* ----------------------------------------------------------------------- *
 * Learned planner (no counterpart in the reference, which has only the
 * analytic model of adatm_bench.c:108-142): an MLP 24 -> 64 -> 64 -> 64 -> 1
 * (ReLU, FP32) predicting log10(ms) of a plan's MTTKRP sequence on the GPU
 * from adatm_plan_features(); trained and evaluated by HIP kernels.
 * ----------------------------------------------------------------------- */
#define ADATM_MLP_IN 24
#define ADATM_MLP_HID 64

typedef struct adatm_mlp adatm_mlp;

size_t adatm_mlp_nparams(void); /* 9985: W1 b1 W2 b2 W3 b3 W4 b4, W row-major in x out */
int adatm_mlp_create(int device, uint64_t seed, adatm_mlp **out);
void adatm_mlp_destroy(adatm_mlp *m);
int adatm_mlp_set_params(adatm_mlp *m, float const *p, size_t n); /* resets Adam */
int adatm_mlp_get_params(adatm_mlp *m, float *p, size_t n);
/* input standardisation x' = (x - xmean) / xstd, target y' = (y - ymean) / ystd */
int adatm_mlp_set_norm(adatm_mlp *m, float const *xmean, float const *xstd,
                       float ymean, float ystd);
int adatm_mlp_get_norm(adatm_mlp *m, float *xmean, float *xstd, float *ymean,
                       float *ystd);
/* mean squared error of the standardised targets over n rows and its
 * gradient w.r.t. the parameters (nparams floats) */
int adatm_mlp_loss_grad(adatm_mlp *m, float const *X, float const *y, uint32_t n,
                        float *grad, float *loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch (X, y) */
int adatm_mlp_step(adatm_mlp *m, float const *X, float const *y, uint32_t n,
                   float lr, float *loss);
/* Adam on shuffled minibatches; sets the standardisation from (X, y) first.
 * loss_hist (nullable, epochs entries) receives the mean batch loss. */
int adatm_mlp_train(adatm_mlp *m, float const *X, float const *y, uint32_t n,
                    uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                    float *loss_hist);
int adatm_mlp_predict(adatm_mlp *m, float const *X, uint32_t n, float *y);
/* adatm_mlp_predict without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code around HIP graph
 * capture, tight planning loops).  The activations live in a device array of
 * the library sized for ADATM_NOALLOC_MAX_ROWS rows (6.7 MB, held by every
 * GPU the library runs on until the process ends); larger n is rejected with
 * SPLATT_ERROR_BADINPUT.  Calls on one GPU share that array: overlapping
 * calls (several host threads, one GPU) race; calls in sequence are safe (a
 * call returns after its work finished).  Same results as adatm_mlp_predict. */
#define ADATM_NOALLOC_MAX_ROWS 4096u
int adatm_mlp_predict_noalloc(adatm_mlp *m, float const *X, uint32_t n, float *y);
int adatm_mlp_save(adatm_mlp *m, char const *path);
int adatm_mlp_load(int device, char const *path, adatm_mlp **out);

/* The 24 features of every candidate (nc x 24, row-major); fiber counts on
 * the device when device >= 0. */
int adatm_plan_features(adatm_coo const *t, uint64_t rank,
                        adatm_plan const *cands, int nc, int device,
                        float *feat);
/* Same from known fiber counts: Fg[g * N + l] for every group g of p. */
int adatm_plan_features_from_counts(uint32_t nmodes, uint64_t nnz,
                                    uint64_t const *dims, uint64_t rank,
                                    adatm_plan const *p, uint64_t const *Fg,
                                    float *feat);
/* Every candidate timed on the device (one CP-ALS sweep in a hipGraph at
 * `precision`, median of nrep after nwarm); the chosen one is the fastest.
 * ms (nullable, one per candidate) receives the timings (+inf: not built). */
int adatm_plan_exhaustive(adatm_coo const *t, uint64_t rank, int precision,
                          int device, int nwarm, int nrep, adatm_plan *cands,
                          int maxc, int *ncands, int *chosen, double *ms);
/* Path of the trained cost model ($ADATM_COSTMODEL, else
 * planner/costmodel.bin beside the library's lib/ directory); 0 if none. */
int adatm_costmodel_path(char *buf, size_t n);
/* Candidates as adatm_plan_analytic; the chosen one minimises the model's
 * prediction (pred, nullable, receives them), or is the analytic choice when
 * model is NULL. */
int adatm_plan_learned(adatm_coo const *t, uint64_t rank, adatm_mlp *model,
                       int device, adatm_plan *cands, int maxc, int *ncands,
                       int *chosen, float *pred);

/* ----------------------------------------------------------------------- *
 * Device engine: group CSFs, memo buffers, factors, output on one GPU.
 * ----------------------------------------------------------------------- */
typedef struct adatm_engine adatm_engine;

int adatm_engine_create(adatm_coo const *t, uint64_t rank,
                        adatm_plan const *plan, int precision, int device,
                        adatm_engine **out);
/* The same with the work-item tiles given: bottom children per reduction
 * item and nodes per scatter item (0 = default 128 / 64; range 8..4096). */
int adatm_engine_create_tiled(adatm_coo const *t, uint64_t rank,
                              adatm_plan const *plan, int precision, int device,
                              uint32_t red_tile, uint32_t scat_tile,
                              adatm_engine **out);
/* The same with creation flags for the root MTTKRP of a memo group: by
 * default the leaf pass goes on to the root in one launch (up to 8 levels)
 * and writes the memo levels it crosses beside its output;
 * ADATM_ENGINE_STAGED (every group) or ADATM_ENGINE_STAGED_GROUP(g) (group g)
 * makes it a chain of passes that stop at every memo level (each memo level
 * written by one pass and read back by the next).  Results differ only in
 * summation grouping. */
#define ADATM_ENGINE_STAGED 0x1u
#define ADATM_ENGINE_STAGED_GROUP(g) (0x100u << (g))
/* The fused reductions read intermediate nodes' factor rows by raw buffer
 * loads when every such factor spans less than 2^31 bytes (a child that does
 * not close its node then loads nothing); with this flag they always use
 * the path of larger factors (a stand-in row).  Results are the same bits. */
#define ADATM_ENGINE_GLOBAL_NODE_ROWS 0x2u
int adatm_engine_create_ex(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                           int precision, int device, uint32_t red_tile,
                           uint32_t scat_tile, uint32_t flags, adatm_engine **out);
/* Children per batch of the fused reductions (4, 8, 16; 0 = default 8, or 4
 * for passes fusing more than 3 levels, with leaf passes in the wide layout
 * below where it applies; 16 applies to passes of <= 2 levels, 8 at most to
 * 3; explicit values keep every pass in the narrow layout unless ORed with
 * ADATM_BATCH_WIDE), optionally ORed with ADATM_BATCH_PIPELINED: the next
 * batch's rows are loaded while one is folded (passes of <= 3 levels, batch
 * 4 or 8).  A launch-time choice: no rebuild. */
#define ADATM_BATCH_PIPELINED 0x100u
/* ORed into the batch: each lane holds 16 bytes of a row (4 FP32/BF16 or 2
 * FP64 columns, so R/4 or R/2 lanes per work item) and gathers rows with
 * 128-bit loads; applies when R is a multiple and >= 8 lanes remain, with
 * batch 4 or 8, not pipelined. */
#define ADATM_BATCH_WIDE 0x200u
int adatm_engine_set_batch(adatm_engine *e, uint32_t batch);
/* Autotuner: the tiles, root-pass form (flags for adatm_engine_create_ex)
 * and batch of one plan, chosen by timing CP-ALS sweeps (nrep, median) of
 * engines built with each candidate; a candidate replaces the defaults
 * (timed first) only when it is more than 1% faster.  batch and flags are
 * nullable. */
int adatm_autotune_tiles(adatm_coo const *t, uint64_t rank, adatm_plan const *plan,
                         int precision, int device, int nrep, uint32_t *red_tile,
                         uint32_t *scat_tile, uint32_t *batch, uint32_t *flags,
                         double *best_ms);
/* The autotuner's search over any timer: timer(ctx, red_tile, scat_tile,
 * batch, flags, &ms) times one candidate (batch as adatm_engine_set_batch,
 * flags as adatm_engine_create_ex) and returns SPLATT_SUCCESS; in this
 * order: the reduction tiles {128, 32, 64, 256, 512} (scatter tile 64), the
 * scatter tiles {16, 32, 128, 256} (best reduction tile), the reduction tiles
 * again when the scatter tile moved, then the flags staging every group in
 * the mask `groups` (the memo groups), then for each such group g the best
 * flags so far with g's ADATM_ENGINE_STAGED_GROUP(g) bit flipped, then the
 * batches {4, 8, 16, 4|PIPELINED, 8|PIPELINED, 4|WIDE, 8|WIDE} with
 * the best flags; a candidate wins only when more than 1% faster than the
 * best so far.  adatm_autotune_tiles is this search over device CP-ALS
 * sweeps. */
typedef int (*adatm_tile_timer)(void *ctx, uint32_t red_tile, uint32_t scat_tile,
                                uint32_t batch, uint32_t flags, double *ms);
int adatm_autotune_search(adatm_tile_timer timer, void *ctx, uint32_t groups,
                          uint32_t *red_tile, uint32_t *scat_tile, uint32_t *batch,
                          uint32_t *flags, double *best_ms);
void adatm_engine_destroy(adatm_engine *e);

/* Factor matrices are exchanged in FP64 row-major (dims[m] x rank). */
int adatm_engine_set_factor(adatm_engine *e, uint32_t mode, double const *A);
int adatm_engine_get_factor(adatm_engine *e, uint32_t mode, double *A);
int adatm_engine_get_lambda(adatm_engine *e, double *lambda);
int adatm_engine_init_factors(adatm_engine *e, uint64_t seed);

/* One MTTKRP of the memoized sequence for `mode`.  A root (tag 0) mode
 * recomputes its group's memo; a reuse mode reads the memo left by its
 * group's root, which must have run since the factors it contracted last
 * changed (same contract as the reference).  out (dims[mode] x rank, FP64)
 * may be NULL to leave the result on the device. */
int adatm_engine_mttkrp(adatm_engine *e, uint32_t mode, double *out);

/* All N MTTKRPs in order with the factors held fixed (splatt bench). */
int adatm_engine_mttkrp_seq(adatm_engine *e);

/* CP-ALS: niters sweeps (stop early when |dfit| < tol and it > 0).
 * fits (niters entries, nullable) receives the fit after every sweep.
 * use_graph = 1 captures one sweep in a hipGraph and replays it. */
int adatm_engine_cpd(adatm_engine *e, uint32_t niters, double tol,
                     int use_graph, double *fits, uint32_t *iters_done);

/* Median wall time (ms) over nrep of: what = 0 the MTTKRP sequence,
 * 1 one full ALS sweep, after nwarm untimed repetitions (hipEvents on the
 * engine's stream).  per_kernel (nullable, 24 entries) then receives, from
 * one extra pass with an event pair around every launch, per kernel class
 * c (0 leaf SpTTM, 1 level Hadamard-reduce, 2 reuse scatter, 3 dense ALS,
 * 4 merge of split parents):
 * [c] total device ms, [8+c] launch count, [16+c] algorithmic bytes. */
int adatm_engine_time(adatm_engine *e, int what, int nwarm, int nrep,
                      int use_graph, double *ms_median, double *per_kernel);

/* Launch records of the last instrumented pass of adatm_engine_time (class
 * as above, 4 = merge of split parents; info = levels fused by a reduction or
 * the level of a scatter; device ms; algorithmic bytes).  Returns the number
 * written (<= max). */
int adatm_engine_launches(adatm_engine *e, int max, int *cls, int *info,
                          double *ms, double *bytes);

/* nsteps back-to-back MTTKRP sequences (what = 0) or ALS sweeps (what = 1,
 * after fresh Grams), bracketed by stream synchronisation; *ms_total is the
 * device time between two events around them. */
int adatm_engine_run_timed(adatm_engine *e, int what, uint32_t nsteps,
                           int use_graph, double *ms_total);

/* Stepwise CP-ALS: begin (Grams of the current factors, sweep counter and
 * solve status reset), then one synchronous sweep at a time returning its
 * fit.  `it` must equal the number of sweeps run since the begin (0, 1, ...);
 * at most 4096 sweeps per begin.  Returns SPLATT_ERROR_NUMERIC when a solve
 * found no positive pivot (see adatm_engine_solve_status). */
int adatm_engine_cpd_begin(adatm_engine *e);
int adatm_engine_cpd_sweep(adatm_engine *e, uint32_t it, int use_graph,
                           double *fit);

/* Worst status of the (hadamard of Grams)^-1 solves since the last CP-ALS
 * start: 0 every pivot positive (the reference's unregularised solve);
 * 1 some solve needed a ridge (1e-10, then 1e-6, of the largest diagonal:
 * the normal equations were not numerically positive definite, and the
 * result departs from an unregularised solve); 2 not even then (factors and
 * fit are not finite; adatm_engine_cpd and the sweeps return
 * SPLATT_ERROR_NUMERIC). */
int adatm_engine_solve_status(adatm_engine *e, int *status);

/* Device bytes of a group's memo at a CSF level (0 when not memoised). */
uint64_t adatm_engine_memo_bytes(adatm_engine const *e, uint32_t grp,
                                 uint32_t level);

/* fits of the first n sweeps since the last CP-ALS start */
int adatm_engine_get_fits(adatm_engine *e, uint32_t n, double *fits);

/* Structure statistics: nfibs per level of each group (n_grp x N, row-major),
 * algorithmic (compulsory one-touch) bytes and flops of one MTTKRP
 * sequence, and device bytes held. */
int adatm_engine_stats(adatm_engine *e, uint64_t *nfibs, double *seq_bytes,
                       double *seq_flops, uint64_t *device_bytes);

/* Copy a group's device CSF back (u32 structure, FP64 values) for checks.
 * Arrays sized by adatm_engine_stats' nfibs: fptr[l] has F_l+1 entries for
 * l < N-1, fids[l] F_l entries for 1 <= l < N (fids[0] is the ordinal). */
int adatm_engine_get_csf(adatm_engine *e, uint32_t grp, uint32_t **fptr,
                         uint32_t **fids, double *vals);

/* ----------------------------------------------------------------------- *
 * Batched CP-ALS of many independent engines on one device (no counterpart
 * in the reference): one hipGraph holds every instance's ALS sweep, so one
 * graph launch runs a batched sweep.  When the engines share one precision
 * the sweep runs in rounds: round r forks every instance with more than r
 * modes onto its own stream for its mode-r MTTKRP, then one grouped launch
 * per dense step updates all of them (results equal to each engine's own
 * sweep, bit for bit); otherwise each instance's whole sweep is a branch.
 * The batch borrows the engines (destroy them after the batch).
 * ----------------------------------------------------------------------- */
typedef struct adatm_batch adatm_batch;
int adatm_batch_create(adatm_engine *const *engines, int n, adatm_batch **out);
void adatm_batch_destroy(adatm_batch *b);
/* niters batched sweeps from the current factors; *ms_total = device time;
 * fits (nullable, n x niters) = every instance's fit after every sweep */
int adatm_batch_cpd(adatm_batch *b, uint32_t niters, double *ms_total, double *fits);
/* each instance's niters sweeps (its own graph), one instance after another */
int adatm_batch_sequential(adatm_batch *b, uint32_t niters, double *ms_total);
/* Persistent mode (on = 1; it applies to batches of one precision with
 * 8 <= R <= 32 and at most 6 modes per instance): each round's MTTKRPs of
 * all instances run in ONE persistent kernel launch, whose workgroups take
 * the recorded launches' blocks from per-XCD queues in dependency order,
 * instead of one forked branch of launches per instance; the results are
 * bit-identical.  on = 0 selects the forked branches (see also
 * adatm_batch_set_form).  adatm_batch_persistent: whether the last built
 * batched sweep ran in persistent mode. */
int adatm_batch_set_persistent(adatm_batch *b, int on);
int adatm_batch_persistent(adatm_batch const *b);
/* The form of each round's MTTKRPs: ADATM_FORM_FORKED (one forked branch of
 * launches per instance), ADATM_FORM_PERSISTENT (above), or
 * ADATM_FORM_GROUPED (the default where it applies): the instances' launches
 * recorded as stages and run as ordinary launches, one per depth of the
 * instances' launch sequences and kind of stage, each covering every
 * instance at that depth (same eligibility as the persistent form;
 * bit-identical results; on 64 uneven instances 3.5 ms per sweep against 5.0
 * for the branches and 5.3 for the persistent launch).  A batch outside the
 * eligibility runs the branches.  adatm_batch_form: the form of the last
 * built batched sweep. */
enum { ADATM_FORM_FORKED = 0, ADATM_FORM_PERSISTENT = 1, ADATM_FORM_GROUPED = 2 };
int adatm_batch_set_form(adatm_batch *b, int form);
int adatm_batch_form(adatm_batch const *b);
/* The queue plan of one round of the persistent form: n chains (an
 * instance's recorded launches), chain i of len[i] stages whose costs are
 * stored chain after chain in cost[]; queue[i] = the queue of chain i
 * (largest total cost first onto the least-loaded of nq queues; ties to the
 * lower chain, then the lower queue); order[qstart[q] .. qstart[q+1]) = the
 * stages (indices into the flat arrays) of queue q, depth by depth, the
 * chains of a depth in the order they were assigned, so every stage follows
 * its chain's previous stage in its queue. */
int adatm_persist_plan(uint32_t n, uint32_t const *len, double const *cost, uint32_t nq,
                       uint32_t *queue, uint32_t *order, uint32_t *qstart);
/* the persistent launch's queues (one per XCD id the device reported),
 * xmap[16] = queue of each XCD id (~0: none) and its workgroups */
int adatm_batch_persist_info(adatm_batch const *b, uint32_t *nq, uint32_t *xmap,
                             uint32_t *grid);

/* Last device error string (thread-local), "" when none. */
char const *adatm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
