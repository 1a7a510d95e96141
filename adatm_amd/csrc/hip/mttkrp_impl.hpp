/*
 * Memoized-MTTKRP kernels for MI355X (gfx950, wave64): the templates,
 * instantiated per value type by mttkrp_fp64.hip, mttkrp_fp32.hip and
 * mttkrp_bf16.hip (three translation units that compile in parallel).
 *
 * A lane group of G lanes owns one work item; lane j holds V consecutive
 * columns of every R-wide row (G = R/V: V = 1, or in the wide layout 16
 * bytes per lane -- 4 FP32/BF16 or 2 FP64 columns), so each factor-row
 * gather and each memo-row access of a group is one contiguous R*sizeof(T)
 * transaction, made of 128-bit lane accesses in the wide layout.  R <= 64;
 * 64/G groups per wavefront.  Work items are balanced by the number of children they walk;
 * a parent longer than one item is cut into pieces whose partial rows go to
 * carry slots, summed in a fixed order by k_fix_chunk / k_fix_multi, so results are
 * deterministic and no output needs zero-filling or atomics.
 *
 *   k_reduce<K>  Z_top(p) = sum over the level-`top+K` descendants c of p of
 *                in(c) o U_K[id(c)], folded up level by level with the
 *                factor rows of the intermediate nodes.  K = 1 is the leaf
 *                SpTTM (p_spttm, src/mttkrp.c:891-928) or one Hadamard-reduce
 *                (p_spt2t_add_hada / p_spt2m_add_hada, src/mttkrp.c:828-888);
 *                K > 1 fuses K such passes so that the levels in between --
 *                the reference's temporaries tmp_vals / tmp_vals_2,
 *                src/mttkrp.c:991-1066 -- never reach HBM.  Per-level running
 *                sums live in registers, as p_propagate_up keeps them in
 *                per-level buffers (src/mttkrp.c:241-304).
 *   k_scatter<D> reuse MTTKRP of CSF level D (p_csf_mttkrp_internal_reuse_
 *                adaptive, src/mttkrp.c:1547-1682): out[i] = sum over level-D
 *                nodes n with id i of Y(n) o Z_D(n), Y(n) = product of the
 *                factor rows of n's ancestors.  The reference scatters under
 *                1024 striped locks; here nodes are pre-sorted by output row
 *                with their ancestors' row ids, so every output row is one
 *                segment of a segmented reduction.
 *   k_fix_chunk  merge of split parents: a lane group per chunk of <= 32 carry
 *                rows; k_fix_multi sums the chunk rows of parents with several.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adatm_kern.h"
#include "common.hpp"

#define PF 4 /* scatter terms whose loads are issued together */

namespace {

/* The item's children are taken in windows of R: lane j of the group loads
 * the metadata of window child j (its id, first differing level and value)
 * in one transaction, the window's close masks come from ballots, and the
 * node ids of the intermediate levels from one more per-lane load, so a
 * window costs two dependent round trips instead of two per child.  The
 * children are then folded in batches of RB (4, 8 or 16; adatm_kchain.rb,
 * tuned per plan): a batch's row loads -- its children's factor rows and
 * those of the intermediate nodes they belong to -- are issued together,
 * with the per-child metadata read from the holding lane (ds_bpermute; for
 * R > 32, one group per wave, v_readlane into scalars). */
template <typename T, typename TU, bool LEAF, int K, int RB, bool DB, bool ONE, int V,
          bool SIDE, bool NB = false>
__device__ __forceinline__ void reduce_body(
    uint32_t vb, uint32_t nitems, uint32_t const *__restrict__ istart,
    uint32_t const *__restrict__ iseg, uint32_t const *__restrict__ islot,
    uint32_t const *__restrict__ ianc, adatm_kchain const &ch,
    T const *__restrict__ xin, T *__restrict__ out, T *__restrict__ carry,
    uint32_t R, uint32_t gpw)
{
  /* a group of G = R/V lanes per item, lane c holding columns [cV, cV+V) */
  uint32_t const G = R / V;
  uint32_t const lane = threadIdx.x & 63u;
  /* ONE (G > 32): one group per wave, so the item index, the window and
     every per-child value read from a lane are wave-uniform (scalar) */
  uint32_t grp, k;
  if constexpr (ONE) {
    if (lane >= G) return;
    grp = 0;
    k = __builtin_amdgcn_readfirstlane(vb * (blockDim.x >> 6) +
                                       (threadIdx.x >> 6));
  } else {
    grp = lane / G;
    if (grp >= gpw) return;
    k = (vb * (blockDim.x >> 6) + (threadIdx.x >> 6)) * gpw + grp;
  }
  uint32_t const gb = grp * G; /* the group's first lane */
  uint32_t const col = lane - gb;
  uint32_t const cv = col * V; /* the lane's first column */
  if (k >= nitems) return;
  /* value v of the group's lane idx */
  auto getl = [&](auto v, uint32_t idx) -> decltype(v) {
    if constexpr (ONE)
      return lane_read(v, idx);
    else
      return lane_get(v, gb + idx);
  };

  uint32_t const a = istart[k], b = istart[k + 1];
  uint32_t const slot = islot[k];
  uint32_t const lt = ch.lt;
  uint8_t const *__restrict__ cfd = ch.cfd;
  uint32_t node[K];
  node[0] = iseg[k];
#pragma unroll
  for (int j = 1; j < K; ++j) node[j] = ianc[(size_t)k * (K - 1) + (j - 1)];
  T acc[K][V];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int v = 0; v < V; ++v) acc[j][v] = (T)0;
  uint32_t const *__restrict__ idb = ch.ids[K];
  TU const *__restrict__ Ub = (TU const *)ch.U[K];
  /* NB: buffer descriptors of the intermediate levels' factors */
  __amdgpu_buffer_rsrc_t nrs[K];
  if constexpr (NB) {
#pragma unroll
    for (int j = 1; j < K; ++j) nrs[j] = buf_rsrc(ch.U[j]);
  }
  /* SIDE: the partial-row slots of the nodes this item's ends cut, and the
     item's first node, at every relative level written beside the top */
  uint32_t shd[K], stl[K], nfst[K];
  if constexpr (SIDE) {
#pragma unroll
    for (int j = 1; j < K; ++j) {
      bool const on = ch.sout[j] != nullptr;
      shd[j] = on ? ch.shead[(size_t)k * K + j] : NONE32;
      stl[j] = on ? ch.stail[(size_t)k * K + j] : NONE32;
      nfst[j] = node[j];
    }
  }

  /* window metadata of child w0 + col (past the item's end: its last
     child); the item's last child closes every level */
  auto load_win = [&](uint32_t w0, uint32_t &mid, uint32_t &mf, T &mx) {
    uint32_t c = w0 + col;
    c = (c < b) ? c : b - 1;
    mid = idb[c];
    uint32_t const fb = cfd[c];
    mf = (c + 1 >= b) ? 0u : fb;
    if constexpr (LEAF) mx = xin[c];
  };
  uint32_t wid, wf;
  T wx = (T)0;
  load_win(a, wid, wf, wx);

  for (uint32_t w0 = a; w0 < b; w0 += G) {
    uint32_t const nw = (b - w0 < G) ? b - w0 : G;
    uint64_t const valid = (nw >= 64u) ? ~0ull : ((1ull << nw) - 1ull);
    uint32_t const me = (col < nw) ? col : nw - 1u;
    uint64_t const below = (1ull << me) - 1ull;
    /* bit i of cm[j]: window child i ends its node of relative level j
       (first differing level <= lt + j; nested, deeper levels close too).
       The lane's child belongs to node node[j] + (closes before it): wn[j]
       is that node's output row (j = 0) or id (j >= 1). */
    uint64_t cm[K];
    uint32_t wn[K];
    uint32_t wnode[K]; /* SIDE: node index of the lane's child at level j */
#pragma unroll
    for (int j = 0; j < K; ++j) {
      cm[j] = ((uint64_t)__ballot(wf <= lt + (uint32_t)j) >> gb) & valid;
      uint32_t const n = node[j] + (uint32_t)__popcll(cm[j] & below);
      if (j == 0)
        wn[0] = ch.orow ? ch.orow[n] : n;
      else
        wn[j] = ch.ids[j][n];
      if constexpr (SIDE) wnode[j] = n;
      node[j] += (uint32_t)__popcll(cm[j]);
    }
    uint32_t nid = 0, nf = 0;
    T nx = (T)0;
    if (w0 + G < b) load_win(w0 + G, nid, nf, nx);

    /* a batch's row loads: its children's factor rows and input values or
       rows, and the factor rows of the level-j nodes holding them */
    auto load_rows = [&](uint32_t s0, T (&xv)[RB][V], T (&uv)[RB][V],
                         T (&uc)[K][RB][V]) {
      uint32_t const nb = (nw - s0 < (uint32_t)RB) ? nw - s0 : (uint32_t)RB;
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        uint32_t const src = s0 + ((uint32_t)i < nb ? (uint32_t)i : nb - 1u);
        uint32_t const id = getl(wid, src);
        ldrow(Ub + (id * R + cv), uv[i]);
        if constexpr (LEAF)
          xv[i][0] = getl(wx, src);
        else
          ldrow(xin + ((size_t)(w0 + src) * R + cv), xv[i]);
#pragma unroll
        for (int j = 1; j < K; ++j) {
          /* a level-j node's row is used once, when the node closes (the
             group's close bit), so it is fetched once per node rather than
             once per child, with no branch among the batch's loads: NB
             (every intermediate factor under 2^31 bytes), a child that does
             not close its node loads through a buffer at an offset past the
             records, which returns zeros without a memory access; else it
             loads row 0, whose line stays cached */
          uint32_t const nid_j = getl(wn[j], src);
          bool const cl = (cm[j] >> (s0 + (uint32_t)i)) & 1u;
          if constexpr (NB) {
            ldrow_buf<TU>(nrs[j], cl ? (nid_j * R + cv) * (uint32_t)sizeof(TU) : BUF_NONE,
                          uc[j][i]);
          } else {
            ldrow((TU const *)ch.U[j] + ((cl ? nid_j : 0u) * R + cv), uc[j][i]);
          }
        }
      }
    };
    auto fold = [&](uint32_t s0, T const (&xv)[RB][V], T const (&uv)[RB][V],
                    T const (&uc)[K][RB][V]) {
      uint32_t const nb = (nw - s0 < (uint32_t)RB) ? nw - s0 : (uint32_t)RB;
      uint32_t cb[K];
#pragma unroll
      for (int j = 0; j < K; ++j) cb[j] = (uint32_t)(cm[j] >> s0);
      auto step = [&](int i) {
#pragma unroll
        for (int v = 0; v < V; ++v)
          acc[K - 1][v] += xv[i][LEAF ? 0 : v] * uv[i][v];
#pragma unroll
        for (int j = K - 1; j >= 1; --j) {
          /* fold the level-j node into its parent with its own factor row */
          bool const c = (cb[j] >> i) & 1u;
          if constexpr (SIDE) {
            /* a closing node of a side level: its row, or the partial row
               of a node this item's ends cut */
            if (c && ch.sout[j]) {
              uint32_t const nd = getl(wnode[j], s0 + (uint32_t)i);
              T *dst;
              if (nd == nfst[j] && shd[j] != NONE32)
                dst = (T *)ch.spart[j] + (size_t)shd[j] * R;
              else if (w0 + s0 + (uint32_t)i + 1u == b && stl[j] != NONE32)
                dst = (T *)ch.spart[j] + (size_t)stl[j] * R;
              else
                dst = (T *)ch.sout[j] + (size_t)nd * R;
              /* memo rows are streamed out: nontemporal, so that they do
                 not evict the gathered factor rows (-3% on a 95M-nonzero
                 5-mode leaf-to-root pass) */
              strow_nt<T, V>(dst + cv, acc[j]);
            }
          }
#pragma unroll
          for (int v = 0; v < V; ++v) {
            T const t = acc[j][v] * uc[j][i][v];
            acc[j - 1][v] = c ? acc[j - 1][v] + t : acc[j - 1][v];
            acc[j][v] = c ? (T)0 : acc[j][v];
          }
        }
        if ((cb[0] >> i) & 1u) {
          uint32_t const row = getl(wn[0], s0 + (uint32_t)i);
          T *dst = (slot != NONE32) ? carry + (size_t)slot * R
                                    : out + (size_t)row * R;
          strow<T, V>(dst + cv, acc[0]);
#pragma unroll
          for (int v = 0; v < V; ++v) acc[0][v] = (T)0;
        }
      };
      if (nb == (uint32_t)RB) {
#pragma unroll
        for (int i = 0; i < RB; ++i) step(i);
      } else {
#pragma unroll
        for (int i = 0; i < RB; ++i)
          if ((uint32_t)i < nb) step(i);
      }
    };
    /* DB: the next batch's rows are loaded while the current one folds */
    T xa[RB][V], ua[RB][V], ca[K][RB][V];
    if constexpr (DB) load_rows(0, xa, ua, ca);
    for (uint32_t s0 = 0; s0 < nw; s0 += RB) {
      if constexpr (DB) {
        bool const more = s0 + RB < nw;
        T xb[RB][V], ub[RB][V], cb2[K][RB][V];
        if (more) load_rows(s0 + RB, xb, ub, cb2);
        fold(s0, xa, ua, ca);
        if (more) {
#pragma unroll
          for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int v = 0; v < V; ++v) {
              xa[i][v] = xb[i][v];
              ua[i][v] = ub[i][v];
#pragma unroll
              for (int j = 1; j < K; ++j) ca[j][i][v] = cb2[j][i][v];
            }
        }
      } else {
        load_rows(s0, xa, ua, ca);
        fold(s0, xa, ua, ca);
      }
    }
    wid = nid;
    wf = nf;
    wx = nx;
  }
}

/* one block of the reduction per workgroup of the launch */
template <typename T, typename TU, bool LEAF, int K, int RB, bool DB, bool ONE, int V,
          bool SIDE, bool NB = false>
__global__ __launch_bounds__(256) void k_reduce(
    uint32_t nitems, uint32_t const *__restrict__ istart,
    uint32_t const *__restrict__ iseg, uint32_t const *__restrict__ islot,
    uint32_t const *__restrict__ ianc, adatm_kchain ch,
    T const *__restrict__ xin, T *__restrict__ out, T *__restrict__ carry,
    uint32_t R, uint32_t gpw)
{
  reduce_body<T, TU, LEAF, K, RB, DB, ONE, V, SIDE, NB>(blockIdx.x, nitems, istart, iseg, islot,
                                                    ianc, ch, xin, out, carry, R, gpw);
}

/* Reuse MTTKRP of level D: nodes pre-sorted by output row, taken in batches
 * of SB whose node ids and ancestor row ids are loaded one batch ahead; the
 * batch's Z rows and ancestor factor rows are then gathered together.  A
 * batch in which no output row ends is summed without per-node checks. */
#define SB 8
static_assert(SB >= 2 && (SB & (SB - 1)) == 0, "the batch sum is a pairwise tree");

/* v[O][c] + ... + v[O+N-1][c] as a pairwise tree: ((v0+v1)+(v2+v3))+... */
template <int N, int O, typename T, int V>
__device__ __forceinline__ T tree_sum(T const (&v)[SB][V], int c)
{
  if constexpr (N == 1)
    return v[O][c];
  else
    return tree_sum<N / 2, O>(v, c) + tree_sum<N / 2, O + N / 2>(v, c);
}

template <typename T, typename TU, int D, bool ZS, int V>
__device__ __forceinline__ void scatter_body(
    uint32_t vb, uint32_t nitems, uint32_t const *__restrict__ istart,
    uint32_t const *__restrict__ iseg, uint32_t const *__restrict__ islot,
    uint32_t const *__restrict__ rowptr, adatm_kscat const &sc,
    T const *__restrict__ Z, T *__restrict__ out, T *__restrict__ carry,
    uint32_t R, uint32_t gpw)
{
  /* a group of G = R/V lanes per item, lane c holding columns [cV, cV+V) */
  uint32_t const G = R / V;
  uint32_t const lane = threadIdx.x & 63u;
  uint32_t const grp = lane / G;
  if (grp >= gpw) return;
  uint32_t const cv = (lane - grp * G) * V;
  uint32_t const k =
      (vb * (blockDim.x >> 6) + (threadIdx.x >> 6)) * gpw + grp;
  if (k >= nitems) return;
  T const zero[V] = {};
  uint32_t const a = istart[k], b = istart[k + 1];
  uint32_t const slot = islot[k];
  bool const heavy = slot != NONE32;
  /* light item: whole rows iseg[k] .. iseg[k+1]-1, positions [a, b) */
  uint32_t p = iseg[k];
  uint32_t const pe = heavy ? p : iseg[k + 1];
  uint32_t e = b; /* end of the current output row */
  if (!heavy) {
    e = rowptr[p + 1];
    while (e == a) { /* leading empty rows */
      strow<T, V>(out + ((size_t)p * R + cv), zero);
      if (++p == pe) return;
      e = rowptr[p + 1];
    }
  }
  /* positions are taken in windows of W = a multiple of SB <= G (G if
     G < SB): lane j < W
     loads the node and ancestor row ids of window position j (past the
     item's end: its last position), and each batch of SB reads them from
     the holding lanes, so a window costs one index round trip */
  uint32_t const gb = grp * G;
  uint32_t const col = lane - gb;
  uint32_t const W = (G >= SB) ? (G / SB) * SB : G;
  /* D > 8: one instance bounds the levels 9..16, the depth sc.d runs */
  constexpr bool RT = D > 8;
  uint32_t const dd = RT ? sc.d : (uint32_t)D;
  auto load_win = [&](uint32_t w0, uint32_t &mn, uint32_t (&mr)[D]) {
    uint32_t q = w0 + (col < W ? col : 0u);
    q = (q < b) ? q : b - 1;
    mn = sc.node[q];
#pragma unroll
    for (int l = 0; l < D; ++l) mr[l] = (!RT || (uint32_t)l < dd) ? sc.rid[l][q] : 0u;
  };
  uint32_t wnd, wrr[D];
  load_win(a, wnd, wrr);
  T acc[V];
#pragma unroll
  for (int c = 0; c < V; ++c) acc[c] = (T)0;
  for (uint32_t w0 = a; w0 < b; w0 += W) {
    uint32_t const we = (b - w0 < W) ? b : w0 + W;
    uint32_t nwnd = 0, nwrr[D];
    if (we < b) load_win(we, nwnd, nwrr);
    for (uint32_t c0 = w0; c0 < we; c0 += SB) {
      uint32_t const nb = (we - c0 < (uint32_t)SB) ? we - c0 : (uint32_t)SB;
      uint32_t nd[SB], rr[D][SB];
#pragma unroll
      for (int i = 0; i < SB; ++i) {
        uint32_t const src = gb + (c0 - w0) + ((uint32_t)i < nb ? (uint32_t)i : nb - 1);
        nd[i] = lane_get(wnd, src);
#pragma unroll
        for (int l = 0; l < D; ++l)
          rr[l][i] = (!RT || (uint32_t)l < dd) ? lane_get(wrr[l], src) : 0u;
      }
      T v[SB][V];
#pragma unroll
      for (int i = 0; i < SB; ++i) {
        T y[V], u[V], z[V];
        ldrow((TU const *)sc.U[0] + (rr[0][i] * R + cv), y);
#pragma unroll
        for (int l = 1; l < D; ++l) {
          if (RT && (uint32_t)l >= dd) continue; /* uniform */
          ldrow((TU const *)sc.U[l] + (rr[l][i] * R + cv), u);
#pragma unroll
          for (int c = 0; c < V; ++c) y[c] = y[c] * u[c];
        }
        if constexpr (ZS) {
          T const zs = Z[nd[i]];
#pragma unroll
          for (int c = 0; c < V; ++c) v[i][c] = zs * y[c];
        } else {
          ldrow(Z + ((size_t)nd[i] * R + cv), z);
#pragma unroll
          for (int c = 0; c < V; ++c) v[i][c] = z[c] * y[c];
        }
      }
      if (nb == (uint32_t)SB && (heavy || e >= c0 + SB)) {
        /* no row ends before the batch's last node (a piece has no row end) */
#pragma unroll
        for (int c = 0; c < V; ++c) acc[c] += tree_sum<SB, 0>(v, c);
        if (!heavy && e == c0 + SB) { /* the row ends with the batch */
          strow<T, V>(out + ((size_t)p * R + cv), acc);
#pragma unroll
          for (int c = 0; c < V; ++c) acc[c] = (T)0;
          while (++p < pe) {
            e = rowptr[p + 1];
            if (e != c0 + SB) break;
            strow<T, V>(out + ((size_t)p * R + cv), zero);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < SB; ++i) {
          uint32_t const c = c0 + i;
          if ((uint32_t)i < nb) {
#pragma unroll
            for (int q = 0; q < V; ++q) acc[q] += v[i][q];
            if (!heavy && c + 1 == e) {
              strow<T, V>(out + ((size_t)p * R + cv), acc);
#pragma unroll
              for (int q = 0; q < V; ++q) acc[q] = (T)0;
              /* next non-empty row; empty rows in between get zeros */
              while (++p < pe) {
                e = rowptr[p + 1];
                if (e != c + 1) break;
                strow<T, V>(out + ((size_t)p * R + cv), zero);
              }
            }
          }
        }
      }
    }
    wnd = nwnd;
#pragma unroll
    for (int l = 0; l < D; ++l) wrr[l] = nwrr[l];
  }
  if (heavy) strow<T, V>(carry + ((size_t)slot * R + cv), acc);
}

template <typename T, typename TU, int D, bool ZS, int V>
__global__ __launch_bounds__(256) void k_scatter(
    uint32_t nitems, uint32_t const *__restrict__ istart,
    uint32_t const *__restrict__ iseg, uint32_t const *__restrict__ islot,
    uint32_t const *__restrict__ rowptr, adatm_kscat sc,
    T const *__restrict__ Z, T *__restrict__ out, T *__restrict__ carry,
    uint32_t R, uint32_t gpw)
{
  scatter_body<T, TU, D, ZS, V>(blockIdx.x, nitems, istart, iseg, islot, rowptr, sc, Z,
                                out, carry, R, gpw);
}

/* Chunk c of split parent h = cseg[c]: the sum of its <= ADATM_FIX_CHUNK
 * carry rows, eight independent partial sums combined in a fixed order; the
 * parent's output row when it has one chunk, else chunk row c of carry2.  A
 * lane group of G = R/V lanes per chunk, lane j holding columns [jV, jV+V):
 * every column sees the same additions whatever V is, so the layouts agree
 * bit for bit. */
template <typename T, int V>
__device__ __forceinline__ void fix_chunk_body(
    uint32_t vb, uint32_t nchunk, uint32_t const *__restrict__ cstart,
    uint32_t const *__restrict__ cseg, uint32_t const *__restrict__ hslot,
    uint32_t const *__restrict__ hchunk, uint32_t const *__restrict__ hseg,
    T const *__restrict__ carry, T *__restrict__ carry2, T *__restrict__ out,
    uint32_t R, uint32_t gpw)
{
  uint32_t const G = R / V;
  uint32_t const lane = threadIdx.x & 63u;
  uint32_t const grp = lane / G;
  if (grp >= gpw) return;
  uint32_t const cv = (lane - grp * G) * V;
  uint32_t const c =
      (vb * (blockDim.x >> 6) + (threadIdx.x >> 6)) * gpw + grp;
  if (c >= nchunk) return;
  uint32_t const h = cseg[c];
  uint32_t const lo = cstart[c];
  uint32_t const end = hslot[h + 1];
  uint32_t const hi = (lo + ADATM_FIX_CHUNK < end) ? lo + ADATM_FIX_CHUNK : end;
  T p[8][V];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int v = 0; v < V; ++v) p[i][v] = (T)0;
  T x[8][V];
  uint32_t s = lo;
  for (; s + 8 <= hi; s += 8) {
#pragma unroll
    for (int i = 0; i < 8; ++i) ldrow(carry + ((size_t)(s + i) * R + cv), x[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int v = 0; v < V; ++v) p[i][v] += x[i][v];
  }
#pragma unroll
  for (int i = 0; i < 7; ++i)
    if (s + i < hi) {
      ldrow(carry + ((size_t)(s + i) * R + cv), x[i]);
#pragma unroll
      for (int v = 0; v < V; ++v) p[i][v] += x[i][v];
    }
  T acc[V];
#pragma unroll
  for (int v = 0; v < V; ++v)
    acc[v] = ((p[0][v] + p[1][v]) + (p[2][v] + p[3][v])) +
             ((p[4][v] + p[5][v]) + (p[6][v] + p[7][v]));
  if (hchunk[h + 1] - hchunk[h] == 1)
    strow<T, V>(out + ((size_t)hseg[h] * R + cv), acc);
  else
    strow<T, V>(carry2 + ((size_t)c * R + cv), acc);
}

/* columns per lane of the merges: 16 bytes when R is a multiple */
template <typename T>
constexpr int fix_v() { return sizeof(T) == 8 ? 2 : 4; }

template <typename T>
__global__ __launch_bounds__(256) void k_fix_chunk(
    uint32_t nchunk, uint32_t const *__restrict__ cstart,
    uint32_t const *__restrict__ cseg, uint32_t const *__restrict__ hslot,
    uint32_t const *__restrict__ hchunk, uint32_t const *__restrict__ hseg,
    T const *__restrict__ carry, T *__restrict__ carry2, T *__restrict__ out,
    uint32_t R, uint32_t gpw)
{
  if (R % fix_v<T>() == 0)
    fix_chunk_body<T, fix_v<T>()>(blockIdx.x, nchunk, cstart, cseg, hslot, hchunk, hseg,
                                  carry, carry2, out, R, gpw);
  else
    fix_chunk_body<T, 1>(blockIdx.x, nchunk, cstart, cseg, hslot, hchunk, hseg, carry,
                         carry2, out, R, gpw);
}

/* Split parents with several chunks (one workgroup each): slice q of the
 * workgroup (G = R/V lanes, lane j holding columns [jV, jV+V)) sums the
 * parent's chunk rows q, q + S, q + 2S, ... (S = 256/G slices) into four
 * running sums taken in turn, and the slices are then added in order.  V
 * depends on R only, so every launch form sums in the same order. */
template <typename T, int V>
__device__ __forceinline__ void fix_multi_body(uint32_t vb, uint32_t const *__restrict__ mlist,
                                               uint32_t const *__restrict__ hchunk,
                                               uint32_t const *__restrict__ hseg,
                                               T const *__restrict__ carry2,
                                               T *__restrict__ out, uint32_t R)
{
  __shared__ T red[256 * V];
  uint32_t const h = mlist[vb];
  uint32_t const G = R / V;
  uint32_t const rpi = 256u / G;
  uint32_t const sl = threadIdx.x / G;
  uint32_t const cv = (threadIdx.x - sl * G) * V;
  T a[4][V];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < V; ++v) a[u][v] = (T)0;
  if (sl < rpi) {
    uint32_t const c1 = hchunk[h + 1];
    uint32_t c = hchunk[h] + sl;
    T x[4][V];
    for (; c + 3 * rpi < c1; c += 4 * rpi) {
#pragma unroll
      for (int u = 0; u < 4; ++u) ldrow(carry2 + ((size_t)(c + u * rpi) * R + cv), x[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < V; ++v) a[u][v] += x[u][v];
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (c + u * rpi < c1) {
        ldrow(carry2 + ((size_t)(c + u * rpi) * R + cv), x[u]);
#pragma unroll
        for (int v = 0; v < V; ++v) a[u][v] += x[u][v];
      }
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
    red[threadIdx.x * V + v] = (a[0][v] + a[1][v]) + (a[2][v] + a[3][v]);
  __syncthreads();
  if (threadIdx.x < G) {
    T t[V];
#pragma unroll
    for (int v = 0; v < V; ++v) t[v] = (T)0;
    for (uint32_t q = 0; q < rpi; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) t[v] += red[(q * G + threadIdx.x) * V + v];
    strow<T, V>(out + ((size_t)hseg[h] * R + cv), t);
  }
}

template <typename T>
__device__ __forceinline__ void fix_multi_any(uint32_t vb, uint32_t const *__restrict__ mlist,
                                              uint32_t const *__restrict__ hchunk,
                                              uint32_t const *__restrict__ hseg,
                                              T const *__restrict__ carry2,
                                              T *__restrict__ out, uint32_t R)
{
  if (R % fix_v<T>() == 0)
    fix_multi_body<T, fix_v<T>()>(vb, mlist, hchunk, hseg, carry2, out, R);
  else
    fix_multi_body<T, 1>(vb, mlist, hchunk, hseg, carry2, out, R);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fix_multi(uint32_t const *__restrict__ mlist,
                                                   uint32_t const *__restrict__ hchunk,
                                                   uint32_t const *__restrict__ hseg,
                                                   T const *__restrict__ carry2,
                                                   T *__restrict__ out, uint32_t R)
{
  fix_multi_any<T>(blockIdx.x, mlist, hchunk, hseg, carry2, out, R);
}

/* ---- launchers ----------------------------------------------------------- */
template <typename T>
hipError_t launch_fixup(hipStream_t s, adatm_kwork const *w, void const *carry,
                        void *carry2, void *out, uint32_t R)
{
  if (w->nchunk == 0) return hipSuccess;
  uint32_t const gpw = 64u / ((R % fix_v<T>() == 0) ? R / fix_v<T>() : R);
  uint32_t const per_block = 4u * gpw;
  hipLaunchKernelGGL((k_fix_chunk<T>), dim3((w->nchunk + per_block - 1) / per_block),
                     dim3(256), 0, s, w->nchunk, w->cstart, w->cseg, w->hslot,
                     w->hchunk, w->hseg, (T const *)carry, (T *)carry2, (T *)out, R, gpw);
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess || w->nmulti == 0) return e;
  hipLaunchKernelGGL((k_fix_multi<T>), dim3(w->nmulti), dim3(256), 0, s, w->mlist,
                     w->hchunk, w->hseg, (T const *)carry2, (T *)out, R);
  return hipGetLastError();
}

template <typename T, typename TU, bool LEAF, int K, int RB, bool DB, int V,
          bool SIDE = false, bool NB = false>
void go_reduce_rb(hipStream_t s, adatm_kwork const *w, adatm_kchain const *ch,
                  void const *xin, void *out, void *carry, uint32_t R)
{
  uint32_t const G = R / V;
  uint32_t const gpw = 64u / G;
  uint32_t const per_block = 4u * gpw;
  dim3 const grid((w->nitems + per_block - 1) / per_block);
  if constexpr (V == 1) {
    if (gpw == 1) {
      hipLaunchKernelGGL((k_reduce<T, TU, LEAF, K, RB, DB, true, 1, SIDE>), grid,
                         dim3(256), 0, s, w->nitems, w->start, w->seg, w->slot, w->anc,
                         *ch, (T const *)xin, (T *)out, (T *)carry, R, gpw);
      return;
    }
  }
  hipLaunchKernelGGL((k_reduce<T, TU, LEAF, K, RB, DB, false, V, SIDE, NB>), grid, dim3(256),
                     0, s, w->nitems, w->start, w->seg, w->slot, w->anc, *ch,
                     (T const *)xin, (T *)out, (T *)carry, R, gpw);
}

/* columns per lane of the wide layout: 16-byte rows pieces (8 for BF16
 * storage, V = 4) when R is a multiple and a group keeps >= 8 lanes */
template <typename T>
constexpr int wide_v() { return sizeof(T) == 8 ? 2 : 4; }

template <typename T, typename TU, bool LEAF, int K>
void go_reduce(hipStream_t s, adatm_kwork const *w, adatm_kchain const *ch,
               void const *xin, void *out, void *carry, uint32_t R)
{
  constexpr int VW = wide_v<T>();
  /* the default batch (rb 0) runs leaf passes in the wide layout: their
     factor-row gathers are what the wide loads speed up */
  bool const wide = (ch->wide || (ch->rb == 0 && !ch->db && LEAF)) && R % VW == 0 &&
                    R / VW >= 8;
  uint32_t const G = wide ? R / VW : R;
  /* a batch never spans windows: no more children per batch than G */
  uint32_t const rb = (ch->rb > G) ? (G >= 8 ? 8u : 4u) : ch->rb;
  if constexpr (LEAF && K >= 2) {
    /* leaf passes writing side levels: batches of 4 or 8, not pipelined */
    if (ch->side) {
      bool const r8 = K <= 3 && rb != 4 && G >= 8;
      if (wide) {
        if (r8) {
          if (ch->nbuf)
            go_reduce_rb<T, TU, LEAF, K, 8, false, VW, true, true>(s, w, ch, xin, out, carry, R);
          else
            go_reduce_rb<T, TU, LEAF, K, 8, false, VW, true>(s, w, ch, xin, out, carry, R);
        } else if (ch->nbuf) {
          go_reduce_rb<T, TU, LEAF, K, 4, false, VW, true, true>(s, w, ch, xin, out, carry, R);
        } else {
          go_reduce_rb<T, TU, LEAF, K, 4, false, VW, true>(s, w, ch, xin, out, carry, R);
        }
      } else if (r8) {
        go_reduce_rb<T, TU, LEAF, K, 8, false, 1, true>(s, w, ch, xin, out, carry, R);
      } else {
        go_reduce_rb<T, TU, LEAF, K, 4, false, 1, true>(s, w, ch, xin, out, carry, R);
      }
      return;
    }
  }
  if (wide) {
    /* node rows through buffers where the factors allow it (passes of
       >= 2 levels; one level has no node rows) */
    bool const nbuf = K >= 2 && ch->nbuf;
    if constexpr (K <= 3) {
      if (rb == 4) {
        if (nbuf)
          go_reduce_rb<T, TU, LEAF, K, 4, false, VW, false, true>(s, w, ch, xin, out, carry, R);
        else
          go_reduce_rb<T, TU, LEAF, K, 4, false, VW>(s, w, ch, xin, out, carry, R);
      } else if (nbuf) {
        go_reduce_rb<T, TU, LEAF, K, 8, false, VW, false, true>(s, w, ch, xin, out, carry, R);
      } else {
        go_reduce_rb<T, TU, LEAF, K, 8, false, VW>(s, w, ch, xin, out, carry, R);
      }
    } else if (nbuf) {
      go_reduce_rb<T, TU, LEAF, K, 4, false, VW, false, true>(s, w, ch, xin, out, carry, R);
    } else {
      go_reduce_rb<T, TU, LEAF, K, 4, false, VW>(s, w, ch, xin, out, carry, R);
    }
    return;
  }
  bool const db = ch->db != 0;
  if constexpr (K <= 3) {
    if (db) {
      if (rb == 4)
        go_reduce_rb<T, TU, LEAF, K, 4, true, 1>(s, w, ch, xin, out, carry, R);
      else
        go_reduce_rb<T, TU, LEAF, K, 8, true, 1>(s, w, ch, xin, out, carry, R);
      return;
    }
  }
  if constexpr (K <= 2) {
    if (rb == 16)
      go_reduce_rb<T, TU, LEAF, K, 16, false, 1>(s, w, ch, xin, out, carry, R);
    else if (rb == 4)
      go_reduce_rb<T, TU, LEAF, K, 4, false, 1>(s, w, ch, xin, out, carry, R);
    else
      go_reduce_rb<T, TU, LEAF, K, 8, false, 1>(s, w, ch, xin, out, carry, R);
  } else if constexpr (K == 3) {
    if (rb == 4)
      go_reduce_rb<T, TU, LEAF, K, 4, false, 1>(s, w, ch, xin, out, carry, R);
    else
      go_reduce_rb<T, TU, LEAF, K, 8, false, 1>(s, w, ch, xin, out, carry, R);
  } else {
    go_reduce_rb<T, TU, LEAF, K, 4, false, 1>(s, w, ch, xin, out, carry, R);
  }
}

template <typename T, typename TU, bool LEAF>
hipError_t reduce_k(hipStream_t s, adatm_kwork const *w, adatm_kchain const *ch,
                    void const *xin, void *out, void *carry, void *carry2, uint32_t R)
{
  if (w->nitems > 0) {
    switch (ch->K) {
    case 1: go_reduce<T, TU, LEAF, 1>(s, w, ch, xin, out, carry, R); break;
    case 2: go_reduce<T, TU, LEAF, 2>(s, w, ch, xin, out, carry, R); break;
    case 3: go_reduce<T, TU, LEAF, 3>(s, w, ch, xin, out, carry, R); break;
    case 4: go_reduce<T, TU, LEAF, 4>(s, w, ch, xin, out, carry, R); break;
    case 5: go_reduce<T, TU, LEAF, 5>(s, w, ch, xin, out, carry, R); break;
    case 6: go_reduce<T, TU, LEAF, 6>(s, w, ch, xin, out, carry, R); break;
    case 7: go_reduce<T, TU, LEAF, 7>(s, w, ch, xin, out, carry, R); break;
    default: go_reduce<T, TU, LEAF, 8>(s, w, ch, xin, out, carry, R); break;
    }
    return hipGetLastError();
  }
  return hipSuccess;
}

template <typename T, typename TU, int D, int V>
void go_scatter_v(hipStream_t s, adatm_kwork const *w, uint32_t const *rowptr,
                  adatm_kscat const *sc, void const *Z, int zs, void *out,
                  void *carry, uint32_t R)
{
  uint32_t const gpw = 64u / (R / V);
  uint32_t const per_block = 4u * gpw;
  dim3 const grid((w->nitems + per_block - 1) / per_block);
  if (zs)
    hipLaunchKernelGGL((k_scatter<T, TU, D, true, V>), grid, dim3(256), 0, s,
                       w->nitems, w->start, w->seg, w->slot, rowptr, *sc,
                       (T const *)Z, (T *)out, (T *)carry, R, gpw);
  else
    hipLaunchKernelGGL((k_scatter<T, TU, D, false, V>), grid, dim3(256), 0, s,
                       w->nitems, w->start, w->seg, w->slot, rowptr, *sc,
                       (T const *)Z, (T *)out, (T *)carry, R, gpw);
}

template <typename T, typename TU, int D>
void go_scatter(hipStream_t s, adatm_kwork const *w, uint32_t const *rowptr,
                adatm_kscat const *sc, void const *Z, int zs, void *out,
                void *carry, uint32_t R)
{
  constexpr int VW = wide_v<T>();
  if (sc->wide && R % VW == 0 && R / VW >= 8)
    go_scatter_v<T, TU, D, VW>(s, w, rowptr, sc, Z, zs, out, carry, R);
  else
    go_scatter_v<T, TU, D, 1>(s, w, rowptr, sc, Z, zs, out, carry, R);
}

template <typename T, typename TU>
hipError_t scatter_d(hipStream_t s, adatm_kwork const *w, uint32_t const *rowptr,
                     adatm_kscat const *sc, void const *Z, int zs, void *out,
                     void *carry, void *carry2, uint32_t R)
{
  if (w->nitems > 0) {
    switch (sc->d) {
    case 1: go_scatter<T, TU, 1>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 2: go_scatter<T, TU, 2>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 3: go_scatter<T, TU, 3>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 4: go_scatter<T, TU, 4>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 5: go_scatter<T, TU, 5>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 6: go_scatter<T, TU, 6>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 7: go_scatter<T, TU, 7>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    case 8: go_scatter<T, TU, 8>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    /* deeper levels (tensors of more than 10 modes): one instance, the
       depth a runtime bound */
    default: go_scatter<T, TU, ADATM_KMAXSCAT>(s, w, rowptr, sc, Z, zs, out, carry, R); break;
    }
    return hipGetLastError();
  }
  return hipSuccess;
}


/* ---- persistent launch (adatm_k_persist) ---------------------------------
 * The stage bodies run in the narrow layout (one lane per column, 64/R items
 * per wave) with the default batches of the single launches; every layout
 * and batch gives bit-identical results, so a stage computes what its own
 * launch would. */

/* the stage descriptor at a wave-uniform address, so that its fields are
   read by scalar loads */
__device__ __forceinline__ adatm_kstage const *ps_uniform(adatm_kstage const *p)
{
  uint64_t const v = (uint64_t)(uintptr_t)p;
  uint32_t const lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t const hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (adatm_kstage const *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

template <typename T, typename TU, bool LEAF, int K, bool SIDE>
__device__ __forceinline__ void ps_reduce(adatm_kstage const *st, uint32_t vb)
{
  st = ps_uniform(st);
  vb = __builtin_amdgcn_readfirstlane(vb);
  constexpr int RB = (K <= 3) ? 8 : 4;
  reduce_body<T, TU, LEAF, K, RB, false, false, 1, SIDE>(
      vb, st->w.nitems, st->w.start, st->w.seg, st->w.slot, st->w.anc, st->ch,
      (T const *)st->in, (T *)st->out, (T *)st->carry, st->R, st->gpw);
}

template <typename T, typename TU, int D, bool ZS>
__device__ __forceinline__ void ps_scatter(adatm_kstage const *st, uint32_t vb)
{
  st = ps_uniform(st);
  vb = __builtin_amdgcn_readfirstlane(vb);
  scatter_body<T, TU, D, ZS, 1>(vb, st->w.nitems, st->w.start, st->w.seg, st->w.slot,
                                st->rowptr, st->sc, (T const *)st->in, (T *)st->out,
                                (T *)st->carry, st->R, st->gpw);
}

template <typename T>
__device__ __forceinline__ void ps_fix_chunk(adatm_kstage const *st, uint32_t vb)
{
  st = ps_uniform(st);
  vb = __builtin_amdgcn_readfirstlane(vb);
  fix_chunk_body<T, 1>(vb, st->w.nchunk, st->w.cstart, st->w.cseg, st->w.hslot, st->w.hchunk,
                    st->w.hseg, (T const *)st->carry, (T *)st->carry2, (T *)st->out, st->R,
                    st->gpw);
}

template <typename T>
__device__ __forceinline__ void ps_fix_multi(adatm_kstage const *st, uint32_t vb)
{
  st = ps_uniform(st);
  vb = __builtin_amdgcn_readfirstlane(vb);
  fix_multi_any<T>(vb, st->w.mlist, st->w.hchunk, st->w.hseg, (T const *)st->carry2,
                    (T *)st->out, st->R);
}

__device__ __forceinline__ void ps_zero(adatm_kstage const *st, uint32_t vb)
{
  st = ps_uniform(st);
  uint32_t *p = (uint32_t *)st->out;
  uint64_t const base = (uint64_t)__builtin_amdgcn_readfirstlane(vb) * ADATM_KS_ZERO_WORDS;
  uint64_t const n = st->nwords;
  for (uint32_t i = threadIdx.x; i < ADATM_KS_ZERO_WORDS; i += 256u)
    if (base + i < n) p[base + i] = 0u;
}

template <typename T, typename TU, bool LEAF, bool SIDE>
__device__ __forceinline__ void ps_reduce_k(adatm_kstage const *st, uint32_t K, uint32_t vb)
{
  switch (K) {
  case 1:
    if constexpr (!SIDE) ps_reduce<T, TU, LEAF, 1, SIDE>(st, vb);
    break;
  case 2: ps_reduce<T, TU, LEAF, 2, SIDE>(st, vb); break;
  case 3: ps_reduce<T, TU, LEAF, 3, SIDE>(st, vb); break;
  case 4: ps_reduce<T, TU, LEAF, 4, SIDE>(st, vb); break;
  default: ps_reduce<T, TU, LEAF, 5, SIDE>(st, vb); break;
  }
}

template <typename T, typename TU, bool ZS>
__device__ __forceinline__ void ps_scatter_d(adatm_kstage const *st, uint32_t d, uint32_t vb)
{
  switch (d) {
  case 1: ps_scatter<T, TU, 1, ZS>(st, vb); break;
  case 2: ps_scatter<T, TU, 2, ZS>(st, vb); break;
  case 3: ps_scatter<T, TU, 3, ZS>(st, vb); break;
  case 4: ps_scatter<T, TU, 4, ZS>(st, vb); break;
  default: ps_scatter<T, TU, 5, ZS>(st, vb); break;
  }
}

/* ---- grouped launch (adatm_k_grouped) -----------------------------------
 * One ordinary launch runs one kind of stage over many stages: block b runs
 * block bmap[b] & 0xffffffff of stage bmap[b] >> 32.  Ordering between
 * launches comes from the stream, as for the single launches. */
template <typename T, typename TU, bool LEAF, int K, bool SIDE>
__global__ __launch_bounds__(256) void k_g_reduce(adatm_kstage const *__restrict__ st,
                                                  uint64_t const *__restrict__ bmap)
{
  uint64_t const m = bmap[blockIdx.x];
  ps_reduce<T, TU, LEAF, K, SIDE>(st + (m >> 32), (uint32_t)m);
}

template <typename T, typename TU, int D, bool ZS>
__global__ __launch_bounds__(256) void k_g_scatter(adatm_kstage const *__restrict__ st,
                                                   uint64_t const *__restrict__ bmap)
{
  uint64_t const m = bmap[blockIdx.x];
  ps_scatter<T, TU, D, ZS>(st + (m >> 32), (uint32_t)m);
}

template <typename T>
__global__ __launch_bounds__(256) void k_g_fix_chunk(adatm_kstage const *__restrict__ st,
                                                     uint64_t const *__restrict__ bmap)
{
  uint64_t const m = bmap[blockIdx.x];
  ps_fix_chunk<T>(st + (m >> 32), (uint32_t)m);
}

template <typename T>
__global__ __launch_bounds__(256) void k_g_fix_multi(adatm_kstage const *__restrict__ st,
                                                     uint64_t const *__restrict__ bmap)
{
  uint64_t const m = bmap[blockIdx.x];
  ps_fix_multi<T>(st + (m >> 32), (uint32_t)m);
}

__global__ __launch_bounds__(256) void k_g_zero(adatm_kstage const *__restrict__ st,
                                                uint64_t const *__restrict__ bmap)
{
  uint64_t const m = bmap[blockIdx.x];
  ps_zero(st + (m >> 32), (uint32_t)m);
}

template <typename T, typename TU, bool LEAF, bool SIDE>
void g_reduce_k(hipStream_t s, uint32_t K, adatm_kstage const *st, uint64_t const *bmap,
                uint32_t nb)
{
  dim3 const g(nb), b(256);
  switch (K) {
  case 1:
    if constexpr (!SIDE) hipLaunchKernelGGL((k_g_reduce<T, TU, LEAF, 1, SIDE>), g, b, 0, s, st, bmap);
    break;
  case 2: hipLaunchKernelGGL((k_g_reduce<T, TU, LEAF, 2, SIDE>), g, b, 0, s, st, bmap); break;
  case 3: hipLaunchKernelGGL((k_g_reduce<T, TU, LEAF, 3, SIDE>), g, b, 0, s, st, bmap); break;
  case 4: hipLaunchKernelGGL((k_g_reduce<T, TU, LEAF, 4, SIDE>), g, b, 0, s, st, bmap); break;
  default: hipLaunchKernelGGL((k_g_reduce<T, TU, LEAF, 5, SIDE>), g, b, 0, s, st, bmap); break;
  }
}

template <typename T, typename TU, bool ZS>
void g_scatter_d(hipStream_t s, uint32_t d, adatm_kstage const *st, uint64_t const *bmap,
                 uint32_t nb)
{
  dim3 const g(nb), b(256);
  switch (d) {
  case 1: hipLaunchKernelGGL((k_g_scatter<T, TU, 1, ZS>), g, b, 0, s, st, bmap); break;
  case 2: hipLaunchKernelGGL((k_g_scatter<T, TU, 2, ZS>), g, b, 0, s, st, bmap); break;
  case 3: hipLaunchKernelGGL((k_g_scatter<T, TU, 3, ZS>), g, b, 0, s, st, bmap); break;
  case 4: hipLaunchKernelGGL((k_g_scatter<T, TU, 4, ZS>), g, b, 0, s, st, bmap); break;
  default: hipLaunchKernelGGL((k_g_scatter<T, TU, 5, ZS>), g, b, 0, s, st, bmap); break;
  }
}

/* one grouped launch of the stages' kind (type, leaf, K, side, d, zs as in
   the stage descriptor; every stage of bmap shares it) */
template <typename T, typename TU>
hipError_t grouped_k(hipStream_t s, uint32_t type, uint32_t leaf, uint32_t K, uint32_t side,
                     uint32_t d, uint32_t zs, adatm_kstage const *st, uint64_t const *bmap,
                     uint32_t nb)
{
  if (nb == 0) return hipSuccess;
  switch (type) {
  case KS_REDUCE:
    if (leaf) {
      if (side)
        g_reduce_k<T, TU, true, true>(s, K, st, bmap, nb);
      else
        g_reduce_k<T, TU, true, false>(s, K, st, bmap, nb);
    } else {
      g_reduce_k<T, TU, false, false>(s, K, st, bmap, nb);
    }
    break;
  case KS_SCATTER:
    if (zs)
      g_scatter_d<T, TU, true>(s, d, st, bmap, nb);
    else
      g_scatter_d<T, TU, false>(s, d, st, bmap, nb);
    break;
  case KS_FIXCHUNK:
    hipLaunchKernelGGL((k_g_fix_chunk<T>), dim3(nb), dim3(256), 0, s, st, bmap);
    break;
  case KS_FIXMULTI:
    hipLaunchKernelGGL((k_g_fix_multi<T>), dim3(nb), dim3(256), 0, s, st, bmap);
    break;
  default: hipLaunchKernelGGL(k_g_zero, dim3(nb), dim3(256), 0, s, st, bmap); break;
  }
  return hipGetLastError();
}

/* XCD (L2 domain) of the calling wave: HW_REG_XCC_ID bits 3:0 */
__device__ __forceinline__ uint32_t xcc_id()
{
  return (uint32_t)__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & 15u;
}

template <typename T, typename TU>
__global__ __launch_bounds__(256) void k_persist(
    adatm_kstage const *__restrict__ stages, uint64_t const *__restrict__ tasks,
    uint32_t const *__restrict__ qoff, uint32_t const *__restrict__ xmap, uint32_t nq,
    uint32_t *ctr, uint32_t *stat, uint32_t spin)
{
  __shared__ uint32_t s_t;
  uint32_t const q = xmap[xcc_id()];
  if (q >= nq) return; /* an XCD without a queue: the whole workgroup leaves */
  uint32_t const q0 = qoff[q], qn = qoff[q + 1] - q0;
  uint32_t *const head = ctr + q;
  uint32_t *const done = ctr + nq;
  uint32_t served = 0;
  uint32_t seen = ADATM_KS_NODEP; /* the last dependency this workgroup acquired */
  for (;;) {
    if (threadIdx.x == 0) s_t = atomicAdd(head, 1u);
    __syncthreads();
    uint32_t const t = __builtin_amdgcn_readfirstlane(s_t);
    if (t >= qn) break;
    uint64_t const task = tasks[q0 + t];
    uint32_t const si = __builtin_amdgcn_readfirstlane((uint32_t)(task >> 32));
    uint32_t const vb = __builtin_amdgcn_readfirstlane((uint32_t)task);
    adatm_kstage const *const st = stages + si;
    uint32_t const dep = st->dep;
    /* a run of tasks with one dependency waits and acquires once: nothing
       rewrites that stage's output while its successors run */
    if (dep != ADATM_KS_NODEP && dep != seen) {
      seen = dep;
      /* every block of the dependency done: the counter is polled in L2
         (sc1 loads), a bounded number of times */
      if (threadIdx.x == 0) {
        uint32_t const need = stages[dep].nblocks;
        uint32_t n = 0;
        while (__hip_atomic_load(done + dep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               need) {
          __builtin_amdgcn_s_sleep(2);
          if (++n >= spin) {
            atomicOr(stat + q, 1u);
            break;
          }
        }
      }
      __syncthreads();
      /* agent-scope acquire (buffer_inv sc1): drops this CU's L1, so that
         rows of buffers a chain rewrites (carry, ping-pong temporaries)
         come from L2, where the producer's stores completed; the
         workgroup-scope invalidate (sc0) leaves the L1 as it is */
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    switch (st->type) {
    case KS_REDUCE:
      if (st->leaf) {
        if (st->side)
          ps_reduce_k<T, TU, true, true>(st, st->K, vb);
        else
          ps_reduce_k<T, TU, true, false>(st, st->K, vb);
      } else {
        ps_reduce_k<T, TU, false, false>(st, st->K, vb);
      }
      break;
    case KS_SCATTER:
      if (st->zs)
        ps_scatter_d<T, TU, true>(st, st->d, vb);
      else
        ps_scatter_d<T, TU, false>(st, st->d, vb);
      break;
    case KS_FIXCHUNK: ps_fix_chunk<T>(st, vb); break;
    case KS_FIXMULTI: ps_fix_multi<T>(st, vb); break;
    default: ps_zero(st, vb); break;
    }
    /* this wave's stores have reached L2 before the block is counted */
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(done + si, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ++served;
    }
  }
  if (threadIdx.x == 0 && served) atomicAdd(stat + nq + q, served);
}

} // namespace

/* the launchers of one value type (T = value/accumulation, TU = factor rows) */
#define ADATM_MTTKRP_INSTANCE(NAME, T, TU)                                          \
  extern "C" hipError_t adatm_reduce_##NAME(hipStream_t s, int leaf,               \
                                            adatm_kwork const *w,                  \
                                            adatm_kchain const *ch, void const *xin, \
                                            void *out, void *carry, void *carry2,  \
                                            uint32_t R)                            \
  {                                                                                \
    return leaf ? reduce_k<T, TU, true>(s, w, ch, xin, out, carry,   \
                                                      carry2, R)                   \
                : reduce_k<T, TU, false>(s, w, ch, xin, out, carry,  \
                                                       carry2, R);                 \
  }                                                                                \
  extern "C" hipError_t adatm_scatter_##NAME(hipStream_t s, adatm_kwork const *w,  \
                                             uint32_t const *rowptr,               \
                                             adatm_kscat const *sc, void const *Z, \
                                             int zs, void *out, void *carry,       \
                                             void *carry2, uint32_t R)             \
  {                                                                                \
    return scatter_d<T, TU>(s, w, rowptr, sc, Z, zs, out, carry,     \
                                          carry2, R);                              \
  }                                                                                \
  extern "C" hipError_t adatm_persist_##NAME(                                       \
      hipStream_t s, uint32_t grid, adatm_kstage const *st, uint64_t const *tasks,  \
      uint32_t const *qoff, uint32_t const *xmap, uint32_t nq, uint32_t *ctr,       \
      uint32_t *stat, uint32_t spin)                                                \
  {                                                                                \
    hipLaunchKernelGGL((k_persist<T, TU>), dim3(grid), dim3(256), 0, s, st, tasks,  \
                       qoff, xmap, nq, ctr, stat, spin);                           \
    return hipGetLastError();                                                      \
  }                                                                                \
  extern "C" hipError_t adatm_grouped_##NAME(                                       \
      hipStream_t s, uint32_t type, uint32_t leaf, uint32_t K, uint32_t side,      \
      uint32_t d, uint32_t zs, adatm_kstage const *st, uint64_t const *bmap,      \
      uint32_t nb)                                                                 \
  {                                                                                \
    return grouped_k<T, TU>(s, type, leaf, K, side, d, zs, st, bmap, nb);          \
  }                                                                                \
  extern "C" hipError_t adatm_persist_grid_##NAME(int device, uint32_t *grid)      \
  {                                                                                \
    int nb = 0, ncu = 0;                                                           \
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                   \
        &nb, k_persist<T, TU>, 256, 0);                                            \
    if (e == hipSuccess)                                                           \
      e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,       \
                                device);                                           \
    *grid = (uint32_t)(nb > 0 ? nb : 1) * (uint32_t)(ncu > 0 ? ncu : 1);           \
    return e;                                                                      \
  }
