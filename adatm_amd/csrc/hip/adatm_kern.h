/*
 * C ABI of the HIP kernels of libsplatt_adatm (launchers, one per kernel).
 * Internal to the library: the host C engine calls these; users call the
 * SPLATT / adatm_* entry points of include/.
 */
#ifndef ADATM_KERN_H
#define ADATM_KERN_H

#include <stdint.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* value layouts of the device path */
enum { KP_FP64 = 0, KP_FP32 = 1, KP_BF16ROWS = 2 };

#define ADATM_KMAXL 16
#define ADATM_KMAXFUSE 8 /* levels one k_reduce launch folds */
#define ADATM_KMAXSCAT 16 /* deepest level k_scatter handles (the engine's mode limit) */

/* Work list of a segmented reduction (see adatm_segwork in the host code):
 * item k walks bottom positions [start[k], start[k+1]); a light item covers
 * whole top parents seg[k] .. seg[k+1]-1, a heavy piece (slot[k] != ~0)
 * part of parent seg[k] and writes carry row slot[k].  anc[k*(K-1)+j-1] is
 * the level-j ancestor (relative to the top) of the item's first position. */
typedef struct {
  uint32_t nitems;
  uint32_t nheavy;
  uint32_t const *start; /* nitems+1 */
  uint32_t const *seg;   /* nitems+1 */
  uint32_t const *slot;  /* nitems */
  uint32_t const *anc;   /* nitems*(K-1), NULL when K == 1 */
  uint32_t const *hseg;  /* nheavy */
  uint32_t const *hslot; /* nheavy+1 */
  /* merge of split parents: chunk c sums carry slots [cstart[c], min(cstart[c]
   * + ADATM_FIX_CHUNK, hslot[cseg[c]+1])); parent h owns chunks [hchunk[h],
   * hchunk[h+1]); the nmulti parents of mlist own more than one chunk */
  uint32_t nchunk, nmulti;
  uint32_t const *cstart; /* nchunk */
  uint32_t const *cseg;   /* nchunk */
  uint32_t const *hchunk; /* nheavy+1 */
  uint32_t const *mlist;  /* nmulti */
} adatm_kwork;

#ifndef ADATM_FIX_CHUNK
#define ADATM_FIX_CHUNK 32u /* carry rows per merge chunk */
#endif

/* K fused CSF levels below a top level lt (relative level 0): ids[j] / U[j]
 * (1 <= j <= K) = node ids and factor of relative level j; cfd[c] = the
 * first (absolute) level at which bottom node c and c+1 have different
 * ancestors (0 for the last node), so the nodes of levels >= cfd[c] end
 * with child c. */
typedef struct {
  uint32_t K, lt;
  uint32_t rb;          /* children per batch: 4, 8 or 16 (0: 8, or 4 if K > 3) */
  uint32_t db;          /* nonzero: load the next batch while one folds (K <= 3) */
  uint32_t wide;        /* nonzero: 16-byte row pieces per lane (R/4 or R/2 lanes
                           per item; R a multiple, >= 8 lanes) */
  uint32_t const *orow; /* output row of each top node (NULL: the node) */
  /* nonzero: every U[j], 1 <= j < K, spans less than 2^31 bytes, so the
     wide-layout reductions read the node rows by raw buffer loads whose
     offsets fit in 32 bits (a child that does not close its node then loads
     nothing) */
  uint32_t nbuf;
  uint8_t const *cfd;
  uint32_t const *ids[ADATM_KMAXFUSE + 1];
  void const *U[ADATM_KMAXFUSE + 1];
  /* side outputs (leaf passes): sout[j] != NULL makes the pass also write
   * the rows of relative level j (the running sum of each node when it
   * closes, before the fold with its own factor row) into sout[j] by node
   * index.  A node that the item's ends cut writes its partial row to
   * spart[j] at slot shead[k*K+j] (the item's first node, cut before the
   * item) or stail[k*K+j] (its last node, cut after it); ~0 = not cut. */
  uint32_t side;
  void *sout[ADATM_KMAXFUSE + 1];
  void *spart[ADATM_KMAXFUSE + 1];
  uint32_t const *shead, *stail;
} adatm_kchain;

/* Z_top(p) = sum over bottom descendants c of p of in(c) o U_K[ids_K[c]],
 * folded level by level with U_j[ids_j[...]] of the intermediate nodes; in(c)
 * is the scalar xin[c] when leaf, else the row xin[c*R ...].  R <= 64. */
int adatm_k_reduce(hipStream_t s, int prec, int leaf, adatm_kwork const *w,
                   adatm_kchain const *ch, void const *xin, void *out,
                   void *carry, void *carry2, uint32_t R);

/* Reuse MTTKRP at CSF level d over level-d nodes sorted by output row:
 * out[i] = sum_{q in [rowptr[i], rowptr[i+1])} Y_q o Z[node[q]],
 * Y_q = U[0][rid[0][q]] o ... o U[d-1][rid[d-1][q]].  With zscalar, Z holds
 * one value per node (the leaf level: Z = the nonzero values). */
typedef struct {
  uint32_t d;
  uint32_t wide; /* nonzero: 16-byte row pieces per lane, as adatm_kchain.wide */
  uint32_t const *node;
  uint32_t const *rid[ADATM_KMAXSCAT];
  void const *U[ADATM_KMAXSCAT];
} adatm_kscat;

int adatm_k_scatter(hipStream_t s, int prec, adatm_kwork const *w,
                    uint32_t const *rowptr, adatm_kscat const *sc,
                    void const *Z, int zscalar, void *out, void *carry,
                    void *carry2, uint32_t R);

/* Merge of the split parents of a work list after adatm_k_reduce /
 * adatm_k_scatter wrote their carry rows (no-op without split parents). */
int adatm_k_fixup(hipStream_t s, int prec, adatm_kwork const *w,
                  void const *carry, void *carry2, void *out, uint32_t R);

/* zero n bytes with a kernel (graph-capture safe) */
int adatm_k_zero(hipStream_t s, void *p, size_t nbytes);

/* ---- persistent launch of many MTTKRP launch sequences ----
 * A stage is one launch of the kernels above (a reduction, a scatter, the two
 * halves of a split-parent merge, or a zero fill), recorded with its
 * arguments instead of launched.  adatm_k_persist runs every stage of many
 * independent launch sequences in one launch: its workgroups take (stage,
 * block) tasks from the queue of the XCD they run on, in queue order, and a
 * task first waits until every block of its stage's `dep` stage is done.
 * Stages that depend on each other are kept on one queue, so the hand-over
 * stays inside one L2: the producer's stores complete (in that L2) before it
 * counts its block done, and the consumer drops its L1 (an agent-scope
 * acquire) before reading.  Every queue is
 * ordered so that a stage's dependency precedes it, so waits end; a wait
 * that still exceeds `spin` polls gives up and flags its queue in err[]. */
enum { KS_ZERO = 0, KS_REDUCE = 1, KS_SCATTER = 2, KS_FIXCHUNK = 3, KS_FIXMULTI = 4 };
#define ADATM_KS_NODEP 0xFFFFFFFFu
#define ADATM_KS_ZERO_WORDS 4096u /* 32-bit words one zero-fill task clears */
typedef struct {
  uint32_t type, nblocks, dep, R, gpw;
  uint32_t leaf, K, side;  /* reduction: leaf pass, levels fused, side outputs */
  uint32_t d, zs;          /* scatter: level, scalar Z */
  uint64_t nwords;         /* zero fill: 32-bit words at out */
  adatm_kwork w;
  adatm_kchain ch;
  adatm_kscat sc;
  uint32_t const *rowptr;
  void const *in;
  void *out, *carry, *carry2;
} adatm_kstage;

/* The persistent launch's layout (the narrow one: one lane per column) needs
 * 8 <= R <= 32; adatm_k_persist_grid = its workgroups for the device. */
int adatm_k_persist_grid(int prec, int device, uint32_t *grid);
/* tasks[qoff[q] .. qoff[q+1]) = queue q, each (stage << 32 | block); xmap[x]
 * = the queue of XCD x (x < 16; ~0: its workgroups take none); ctr =
 * nq heads then one done counter per stage, all zero at launch; stat = nq
 * error flags then nq served-task counts, accumulated over launches. */
int adatm_k_persist(hipStream_t s, int prec, uint32_t grid, adatm_kstage const *stages,
                    uint64_t const *tasks, uint32_t const *qoff, uint32_t const *xmap,
                    uint32_t nq, uint32_t *ctr, uint32_t *stat, uint32_t spin);
/* Grouped launch: one ordinary launch of nblocks blocks runs the stages of
 * one kind (kind->type, leaf, K, side, d, zs; K and d <= 5): block b runs
 * block bmap[b] & 0xffffffff of stages[bmap[b] >> 32]. */
int adatm_k_grouped(hipStream_t s, int prec, adatm_kstage const *kind,
                    adatm_kstage const *stages, uint64_t const *bmap, uint32_t nblocks);
/* XCD ids (HW_REG_XCC_ID) seen by nblocks workgroups: out[b] for block b */
int adatm_k_xcc_probe(hipStream_t s, uint32_t *out, uint32_t nblocks);

/* Dense CP-ALS steps (FP64 arithmetic; factors/M in the path's precision).
 * G: N grams of R x R (FP64) contiguous.  part: scratch of
 * adatm_k_part_doubles(R) doubles. */
size_t adatm_k_part_doubles(uint32_t R);
int adatm_k_gram(hipStream_t s, int prec, void const *A, uint64_t I,
                 uint32_t R, double *part, double *G_out);
/* sum of squares of n values (FP64 accumulation) into *out */
int adatm_k_sqnorm(hipStream_t s, int prec, void const *x, uint64_t n,
                   double *part, double *out);
int adatm_k_solve_inv(hipStream_t s, double const *G, uint32_t N, uint32_t m,
                      uint32_t R, double *Vinv, int *status);
/* A = M Vinv, then lambda (2-norm when *iter == 0, else max(1, colmax)),
 * A /= lambda, G_out = A^T A, bf16 copy refreshed when Abf != NULL.
 * `iter` is a device counter so that one captured graph serves all sweeps. */
int adatm_k_update(hipStream_t s, int prec, void const *M, double const *Vinv,
                   void *A, void *Abf, uint64_t I, uint32_t R,
                   uint32_t const *iter, double *lambda, double *part,
                   double *G_out);
/* fit of the current model into fits[*iter] (skipped when *iter >= fcap,
 * the entries of fits), then ++*iter (needs M of the last mode). */
int adatm_k_fit(hipStream_t s, int prec, void const *A_last, void const *M,
                uint64_t I_last, double const *lambda, double const *G,
                uint32_t N, uint32_t R, double ttnormsq, double *part,
                double *fits, uint32_t fcap, uint32_t *iter);
/* The dense update of many CP-ALS instances at once (one per descriptor,
 * each at its own mode m): solve, A = M Vinv with column statistics, lambda,
 * normalisation and Gram, with every instance's arithmetic, block split and
 * summation order those of adatm_k_solve_inv + adatm_k_update on it alone.
 * Descriptors and block maps live in device memory; bm_* map every block of
 * a grouped grid to its descriptor (adatm_k_dense_blocks gives the counts). */
typedef struct {
  uint32_t N, m, R;
  uint32_t fcap;         /* fit: entries of fits */
  uint64_t I;
  double const *G;       /* N grams R x R */
  double *Vinv;
  int *status;
  void const *M;
  void *A, *Abf;
  uint32_t *iter;
  double *lambda, *part, *G_out;
  uint32_t mm0, nmm;     /* first block and block count of the product */
  uint32_t gr0, ngr;     /* ... of the Gram */
  uint32_t sp0, nsp;     /* ... of the Gram's partial sums */
  uint32_t in0, nin;     /* fit: first block and block count of the inner product */
  double ttnormsq;       /* fit: squared norm of the tensor */
  double *fits;          /* fit: fits[*iter] is written, then ++*iter */
} adatm_kdense;
void adatm_k_dense_blocks(uint64_t I, uint32_t R, uint32_t *nmm, uint32_t *ngr,
                          uint32_t *nsp);
int adatm_k_dense_grouped(hipStream_t s, int prec, adatm_kdense const *desc, uint32_t n,
                          uint32_t Rmax, uint32_t const *bm_mm, uint32_t nb_mm,
                          uint32_t const *bm_gr, uint32_t nb_gr, uint32_t const *bm_sp,
                          uint32_t nb_sp);
/* the fits of many instances (descriptors with m = N-1: A the last factor,
 * M its MTTKRP, I its rows; in0/nin, ttnormsq, fits, iter), as adatm_k_fit
 * on each; adatm_k_fit_blocks(I) = the inner product's block count */
uint32_t adatm_k_fit_blocks(uint64_t I);
int adatm_k_fit_grouped(hipStream_t s, int prec, adatm_kdense const *desc, uint32_t n,
                        uint32_t const *bm_in, uint32_t nb_in);
/* bits of the BF16 rounding (nearest even; NaN stays a quiet NaN) that
 * the BF16 factor-row copies use, evaluated on the host */
uint16_t adatm_k_bf16_bits(float f);
/* FP64 host matrix <-> device factor in the path's precision */
int adatm_k_convert_in(hipStream_t s, int prec, double const *src64,
                       void *dst, void *dst_bf, uint64_t n);
int adatm_k_convert_out(hipStream_t s, int prec, void const *src,
                        double *dst64, uint64_t n);

/* ---- construction (construct.hip) ---- */
uint64_t adatm_k_scan_temp_words(uint64_t n);
int adatm_k_exscan(hipStream_t s, uint32_t const *in, uint32_t *out, uint64_t n,
                   uint32_t *temp, uint32_t *total);
uint64_t adatm_k_sort_temp_words(uint64_t n);
int adatm_k_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint64_t n,
                       uint32_t nbits, uint32_t *temp);
int adatm_k_iota(hipStream_t s, uint32_t *p, uint64_t n);
int adatm_k_gather(hipStream_t s, uint32_t *dst, uint32_t const *src,
                   uint32_t const *idx, uint64_t n);
int adatm_k_gather_vals(hipStream_t s, int prec, void *dst, double const *src,
                        uint32_t const *idx, uint64_t n);
int adatm_k_first_diff(hipStream_t s, uint32_t const *const *sc, uint32_t N,
                       uint64_t n, uint8_t *fd);
int adatm_k_level_flags(hipStream_t s, uint8_t const *fd, uint64_t n, uint32_t l,
                        uint32_t *flags);
int adatm_k_level_emit(hipStream_t s, uint32_t const *flags, uint32_t const *pos,
                       uint32_t const *scl, uint64_t n, uint32_t *fids,
                       uint32_t *start);
int adatm_k_fptr(hipStream_t s, uint32_t const *start, uint64_t F,
                 uint32_t const *pos_next, uint32_t F_next, uint32_t *fptr);
int adatm_k_parent(hipStream_t s, uint32_t const *start_next, uint64_t Fn,
                   uint32_t const *flags, uint32_t const *pos, uint32_t *par);
int adatm_k_compose(hipStream_t s, uint32_t *cp, uint64_t n1, uint32_t const *fp);
/* cfd of level lb from the parent arrays par[1..lb] (par[l]: level l -> l-1) */
int adatm_k_cfd(hipStream_t s, uint32_t const *const *par, uint32_t lb, uint64_t F,
                uint8_t *cfd);
int adatm_k_items_count(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                        uint32_t tile, uint32_t *nitem, uint32_t *npiece,
                        uint32_t *heavy);
int adatm_k_items_emit(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                       uint32_t tile, uint32_t const *nitem,
                       uint32_t const *ibase, uint32_t const *sbase,
                       uint32_t const *heavy, uint32_t const *hbase,
                       uint32_t *start, uint32_t *seg, uint32_t *slot,
                       uint32_t *hseg, uint32_t *hslot);
int adatm_k_items_tail(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                       uint32_t nitems, uint32_t nheavy, uint32_t nslots,
                       uint32_t *start, uint32_t *seg, uint32_t *slot,
                       uint32_t *hslot);
int adatm_k_fix_count(hipStream_t s, uint32_t const *hslot, uint32_t nheavy,
                      uint32_t ch, uint32_t *nch, uint32_t *multi);
int adatm_k_fix_emit(hipStream_t s, uint32_t const *hslot, uint32_t nheavy,
                     uint32_t ch, uint32_t *hchunk, uint32_t nchunk,
                     uint32_t const *multi, uint32_t const *mbase, uint32_t *cstart,
                     uint32_t *cseg, uint32_t *mlist);
int adatm_k_items_anc(hipStream_t s, uint32_t const *start, uint32_t nitems,
                      uint32_t lb, uint32_t lt, uint32_t const *const *par,
                      uint32_t *anc);
/* side-level cuts of a work list (see k_side_count / k_side_emit): per item
 * the partial rows of the level-L nodes its ends cut, their slots in item
 * order, and the runs of slots that belong to one node */
int adatm_k_side_count(hipStream_t s, uint32_t const *start, uint32_t nitems,
                       uint32_t lb, uint32_t L, uint64_t Fb, uint8_t const *cfd,
                       uint32_t const *const *par, uint32_t *cnt, uint32_t *nf,
                       uint32_t *nl, uint32_t *fl);
int adatm_k_side_emit(hipStream_t s, uint32_t nitems, uint32_t K, uint32_t j,
                      uint32_t const *cnt, uint32_t const *base, uint32_t const *nf,
                      uint32_t const *nl, uint32_t const *fl, uint32_t *head,
                      uint32_t *tail, uint32_t *snode);
int adatm_k_run_flags(hipStream_t s, uint32_t const *v, uint64_t n, uint32_t *flags);
int adatm_k_run_emit(hipStream_t s, uint32_t const *flags, uint32_t const *pos,
                     uint32_t const *v, uint64_t n, uint32_t nrun, uint32_t *val,
                     uint32_t *first);
int adatm_k_rowptr(hipStream_t s, uint32_t const *skey, uint64_t n,
                   uint32_t nrows, uint32_t *rowptr);
int adatm_k_scatter_rids(hipStream_t s, uint32_t const *node, uint64_t n,
                         uint32_t d, uint32_t const *const *par,
                         uint32_t const *const *fids, uint32_t *const *rid);

/* learned cost model network (mlp.hip): FP32, W row-major I x O */
int adatm_k_dense_fwd(hipStream_t s, float const *H, float const *W,
                      float const *b, float *Z, float *Hout, uint32_t B,
                      uint32_t I, uint32_t O, int relu);
int adatm_k_mse(hipStream_t s, float const *y, float const *t, uint32_t B,
                float *dy, float *part);
int adatm_k_dense_bwd(hipStream_t s, float const *H, float const *dZ,
                      float const *W, float const *Zprev, float *dH, float *dW,
                      float *db, uint32_t B, uint32_t I, uint32_t O);
int adatm_k_adam(hipStream_t s, float *p, float const *g, float *m, float *v,
                 uint32_t n, float lr, float b1, float b2, float eps, float c1,
                 float c2);

/* host-side pieces of the cost-model network shared with the allocation-free
 * prediction (learned.c): the forward pass over caller-placed device
 * buffers, and the network's device, stream and standardisation */
struct adatm_mlp;
int adatm_mlp_forward_dev(struct adatm_mlp *M, float const *x, float *const *Z,
                          float *const *H, uint32_t B);
void adatm_mlp_dev_info(struct adatm_mlp const *M, int *device, hipStream_t *stream,
                        float *xmean, float *xstd, float *ymean, float *ystd);

#ifdef __cplusplus
}
#endif

#endif
