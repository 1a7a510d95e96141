/*
 * CSF construction on the GPU (gfx950, wave64).
 *
 *   stable LSD radix sort of (key, value) u32 pairs, 8-bit digits:
 *     k_rs_hist      per-tile digit histogram (LDS atomics)
 *     exclusive scan of the digit-major histogram -> tile offsets
 *     k_rs_scatter   stable rank of each element among equal digits of its
 *                    tile: wave-level match by 8 ballots, per-wave running
 *                    counts in LDS, prefix over the tile's waves
 *   lexicographic order of coordinates by a level order = one key pass per
 *   mode, least significant first (SPLATT's tt_sort, called at
 *   src/adatm_rcsf.c:179; unique result for unique coordinates, and stable
 *   for duplicates);
 *   3-phase exclusive scan of u32;
 *   CSF levels from the sorted coordinates (p_mk_outerptr / p_mk_fptr,
 *   src/adatm_rcsf.c:21-163): a node starts wherever the prefix up to its
 *   level changes (first differing level <= l), node ids by scan, child
 *   pointers and parents by lookups into the next/previous level's scan;
 *   work lists of the segmented reductions (item starts, split pieces,
 *   carry slots, ancestors of item starts) by counting + scan.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adatm_kern.h"
#include "common.hpp"

#define CB 256           /* threads per construction block */
#define RS_ITEMS 16      /* elements per thread per radix tile */
#define RS_TILE (CB * RS_ITEMS)
#define SC_ITEMS 8
#define SC_TILE (CB * SC_ITEMS)

namespace {

inline uint32_t nblk(uint64_t n, uint32_t per)
{
  return (uint32_t)((n + per - 1) / per);
}

/* ---- scan ----------------------------------------------------------------- */
__global__ __launch_bounds__(CB) void k_scan_tile(uint32_t const *__restrict__ in,
                                                  uint32_t *__restrict__ out,
                                                  uint64_t n,
                                                  uint32_t *__restrict__ bsum)
{
  __shared__ uint32_t sh[CB];
  uint64_t const base = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    uint64_t const j = base + i;
    v[i] = (j < n) ? in[j] : 0u;
    s += v[i];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  /* Hillis-Steele inclusive scan of the thread sums */
  for (uint32_t off = 1; off < CB; off <<= 1) {
    uint32_t const t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - s;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    uint64_t const j = base + i;
    if (j < n) out[j] = run;
    run += v[i];
  }
  if (threadIdx.x == CB - 1) bsum[blockIdx.x] = sh[CB - 1];
}

__global__ __launch_bounds__(CB) void k_scan_add(uint32_t *__restrict__ out,
                                                 uint64_t n,
                                                 uint32_t const *__restrict__ boff)
{
  uint64_t const base = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
  uint32_t const add = boff[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    uint64_t const j = base + i;
    if (j < n) out[j] += add;
  }
}

/* total = out[n-1] + in[n-1] into *tot */
__global__ void k_scan_total(uint32_t const *in, uint32_t const *out, uint64_t n,
                             uint32_t *tot)
{
  *tot = (n > 0) ? out[n - 1] + in[n - 1] : 0u;
}

/* exclusive scan; temp holds the block sums of every recursion level */
hipError_t exscan(hipStream_t s, uint32_t const *in, uint32_t *out, uint64_t n,
                  uint32_t *temp)
{
  if (n == 0) return hipSuccess;
  uint32_t const nb = nblk(n, SC_TILE);
  uint32_t *bsum = temp;
  hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(CB), 0, s, in, out, n, bsum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nb == 1) return e;
  uint32_t *boff = temp + nb;
  e = exscan(s, bsum, boff, nb, temp + 2 * (uint64_t)nb);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(CB), 0, s, out, n, boff);
  return hipGetLastError();
}

uint64_t exscan_temp_words(uint64_t n)
{
  uint64_t w = 0;
  while (n > 1) {
    uint64_t const nb = (n + SC_TILE - 1) / SC_TILE;
    w += 2 * nb;
    n = nb;
  }
  return w + 4;
}

/* ---- radix sort ------------------------------------------------------------ */
__global__ __launch_bounds__(CB) void k_rs_hist(uint32_t const *__restrict__ keys,
                                                uint64_t n, uint32_t shift,
                                                uint32_t nb,
                                                uint32_t *__restrict__ ghist)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t const base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    uint64_t const j = base + (uint64_t)i * CB + threadIdx.x;
    if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);
  }
  __syncthreads();
  ghist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

/* stable scatter: tile elements are ordered (wave, step, lane) and each wave
 * owns a contiguous sub-tile, so ranks by (wave prefix, running count,
 * lanes below with the same digit) preserve the input order */
__global__ __launch_bounds__(CB) void k_rs_scatter(
    uint32_t const *__restrict__ kin, uint32_t const *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint64_t n,
    uint32_t shift, uint32_t nb, uint32_t const *__restrict__ goff)
{
  __shared__ uint32_t cnt[4][256];
  __shared__ uint32_t toff[256];
  uint32_t const lane = threadIdx.x & 63u;
  uint32_t const w = threadIdx.x >> 6;
  for (uint32_t x = threadIdx.x; x < 4 * 256; x += CB) (&cnt[0][0])[x] = 0;
  toff[threadIdx.x] = goff[(uint64_t)threadIdx.x * nb + blockIdx.x];
  __syncthreads();
  uint64_t const base = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)w * (RS_ITEMS * 64);
  uint64_t const lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
  for (int s = 0; s < RS_ITEMS; ++s) {
    uint64_t const j = base + (uint64_t)s * 64 + lane;
    bool const ok = j < n;
    uint32_t const k = ok ? kin[j] : 0u;
    key[s] = k;
    val[s] = ok ? vin[j] : 0u;
    uint32_t const d = (k >> shift) & 255u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      uint64_t const b = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? b : ~b;
    }
    uint32_t const before = cnt[w][d];
    rank[s] = ok ? before + (uint32_t)__popcll(peers & lt) : 0u;
    /* every lane of the wave has read its digit's count before the group's
       lowest lane advances it (lockstep order, kept by the fence) */
    __builtin_amdgcn_wave_barrier();
    if (ok && (peers & lt) == 0ull) cnt[w][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  /* prefix of the per-wave counts for this thread's digit */
  {
    uint32_t const d = threadIdx.x;
    uint32_t run = 0;
    for (uint32_t ww = 0; ww < 4; ++ww) {
      uint32_t const c = cnt[ww][d];
      cnt[ww][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < RS_ITEMS; ++s) {
    uint64_t const j = base + (uint64_t)s * 64 + lane;
    if (j < n) {
      uint32_t const d = (key[s] >> shift) & 255u;
      uint32_t const pos = toff[d] + cnt[w][d] + rank[s];
      kout[pos] = key[s];
      vout[pos] = val[s];
    }
  }
}

__global__ void k_iota(uint32_t *p, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = (uint32_t)i;
}

__global__ void k_gather(uint32_t *__restrict__ dst, uint32_t const *__restrict__ src,
                         uint32_t const *__restrict__ idx, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[idx[i]];
}

template <typename T>
__global__ void k_gather_vals(T *__restrict__ dst, double const *__restrict__ src,
                              uint32_t const *__restrict__ idx, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = (T)src[idx[i]];
}

/* ---- CSF levels -------------------------------------------------------- */
struct sc_ptrs {
  uint32_t const *p[ADATM_KMAXL];
};

__global__ void k_first_diff(sc_ptrs sc, uint32_t N, uint64_t n,
                             uint8_t *__restrict__ fd)
{
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += st) {
    uint8_t l = 0;
    if (j > 0) {
      l = (uint8_t)N;
      for (uint32_t k = 0; k < N; ++k)
        if (sc.p[k][j] != sc.p[k][j - 1]) {
          l = (uint8_t)k;
          break;
        }
    }
    fd[j] = l;
  }
}

__global__ void k_level_flags(uint8_t const *__restrict__ fd, uint64_t n,
                              uint32_t l, uint32_t *__restrict__ flags)
{
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += st) flags[j] = (j == 0 || fd[j] <= l) ? 1u : 0u;
}

__global__ void k_level_emit(uint32_t const *__restrict__ flags,
                             uint32_t const *__restrict__ pos,
                             uint32_t const *__restrict__ scl, uint64_t n,
                             uint32_t *__restrict__ fids,
                             uint32_t *__restrict__ start)
{
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += st)
    if (flags[j]) {
      fids[pos[j]] = scl[j];
      start[pos[j]] = (uint32_t)j;
    }
}

/* fptr[p] = node index at the next level of position start[p] (identity for
 * the leaf level); fptr[F] = F_next */
__global__ void k_fptr(uint32_t const *__restrict__ start, uint64_t F,
                       uint32_t const *__restrict__ pos_next, uint32_t F_next,
                       uint32_t *__restrict__ fptr)
{
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; p <= F; p += st)
    fptr[p] = (p == F) ? F_next : (pos_next ? pos_next[start[p]] : start[p]);
}

/* par[c] = node at the parent level containing position start_next[c]
 * (start_next = NULL: the leaf level, position c) */
__global__ void k_parent(uint32_t const *__restrict__ start_next, uint64_t Fn,
                         uint32_t const *__restrict__ flags,
                         uint32_t const *__restrict__ pos,
                         uint32_t *__restrict__ par)
{
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; c < Fn; c += st) {
    uint32_t const s = start_next ? start_next[c] : (uint32_t)c;
    par[c] = pos[s] + flags[s] - 1u;
  }
}

/* ---- work lists ----------------------------------------------------------- */
/* per segment: items it opens, split pieces, heavy flag */
__global__ void k_items_count(uint32_t const *__restrict__ cp, uint64_t nseg,
                              uint32_t tile, uint32_t *__restrict__ nitem,
                              uint32_t *__restrict__ npiece,
                              uint32_t *__restrict__ heavy)
{
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; p < nseg; p += st) {
    uint32_t const len = cp[p + 1] - cp[p];
    if (len > tile) {
      uint32_t const pc = (len + tile - 1) / tile;
      nitem[p] = pc;
      npiece[p] = pc;
      heavy[p] = 1;
    } else {
      bool opens = (p == 0);
      if (!opens) {
        uint32_t const pl = cp[p] - cp[p - 1];
        opens = (pl > tile) || (cp[p] / tile != cp[p - 1] / tile);
      }
      nitem[p] = opens ? 1u : 0u;
      npiece[p] = 0;
      heavy[p] = 0;
    }
  }
}

__global__ void k_items_emit(uint32_t const *__restrict__ cp, uint64_t nseg,
                             uint32_t tile, uint32_t const *__restrict__ nitem,
                             uint32_t const *__restrict__ ibase,
                             uint32_t const *__restrict__ sbase,
                             uint32_t const *__restrict__ heavy,
                             uint32_t const *__restrict__ hbase,
                             uint32_t *__restrict__ start,
                             uint32_t *__restrict__ seg,
                             uint32_t *__restrict__ slot,
                             uint32_t *__restrict__ hseg,
                             uint32_t *__restrict__ hslot)
{
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; p < nseg; p += st) {
    uint32_t const k0 = ibase[p];
    if (heavy[p]) {
      for (uint32_t q = 0; q < nitem[p]; ++q) {
        start[k0 + q] = cp[p] + q * tile;
        seg[k0 + q] = (uint32_t)p;
        slot[k0 + q] = sbase[p] + q;
      }
      hseg[hbase[p]] = (uint32_t)p;
      hslot[hbase[p]] = sbase[p];
    } else if (nitem[p]) {
      start[k0] = cp[p];
      seg[k0] = (uint32_t)p;
      slot[k0] = NONE32;
    }
  }
}

__global__ void k_items_tail(uint32_t const *__restrict__ cp, uint64_t nseg,
                             uint32_t nitems, uint32_t nheavy, uint32_t nslots,
                             uint32_t *start, uint32_t *seg, uint32_t *slot,
                             uint32_t *hslot)
{
  start[nitems] = cp[nseg];
  seg[nitems] = (uint32_t)nseg;
  slot[nitems] = NONE32;
  hslot[nheavy] = nslots;
}

/* split-parent merge lists: chunks of <= ch carry slots per split parent,
 * and the parents spanning more than one chunk */
__global__ void k_fix_count(uint32_t const *__restrict__ hslot, uint32_t nheavy,
                            uint32_t ch, uint32_t *__restrict__ nch,
                            uint32_t *__restrict__ multi)
{
  uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; h < nheavy; h += st) {
    uint32_t const ns = hslot[h + 1] - hslot[h];
    uint32_t const c = (ns + ch - 1) / ch;
    nch[h] = c;
    multi[h] = c > 1 ? 1u : 0u;
  }
}

__global__ void k_fix_emit(uint32_t const *__restrict__ hslot, uint32_t nheavy,
                           uint32_t ch, uint32_t const *__restrict__ hchunk,
                           uint32_t const *__restrict__ multi,
                           uint32_t const *__restrict__ mbase,
                           uint32_t *__restrict__ cstart, uint32_t *__restrict__ cseg,
                           uint32_t *__restrict__ mlist)
{
  uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; h < nheavy; h += st) {
    uint32_t const c0 = hchunk[h];
    uint32_t const c1 = hchunk[h + 1];
    for (uint32_t c = c0; c < c1; ++c) {
      cstart[c] = hslot[h] + (c - c0) * ch;
      cseg[c] = (uint32_t)h;
    }
    if (multi[h]) mlist[mbase[h]] = (uint32_t)h;
  }
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

struct par_ptrs {
  uint32_t const *p[ADATM_KMAXL];
};

/* first level at which nodes c and c+1 of level lb have different ancestors */
__global__ void k_cfd(par_ptrs par, uint32_t lb, uint64_t F, uint8_t *__restrict__ cfd)
{
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; c < F; c += st) {
    uint32_t l = 0;
    if (c + 1 < F) {
      uint32_t x = (uint32_t)c, y = (uint32_t)c + 1;
      l = lb;
      while (l > 0) {
        uint32_t const px = par.p[l][x], py = par.p[l][y];
        if (px == py) break;
        x = px;
        y = py;
        --l;
      }
    }
    cfd[c] = (uint8_t)l;
  }
}

/* anc[k*(K-1) + j-1] = level-(lt+j) ancestor of item k's first position */
__global__ void k_items_anc(uint32_t const *__restrict__ start, uint32_t nitems,
                            uint32_t lb, uint32_t lt, par_ptrs par,
                            uint32_t *__restrict__ anc)
{
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  uint32_t const K = lb - lt;
  for (; k < nitems; k += st) {
    uint32_t v = start[k];
    for (uint32_t l = lb; l > lt + 1; --l) {
      v = par.p[l][v];
      anc[k * (K - 1) + (l - 1 - lt) - 1] = v;
    }
  }
}

/* Nodes of a side level L (lt < L < lb, written by a fused pass beside its
 * top level) cut by the ends of work item k: the item's first node (the
 * level-L ancestor of its first position a) is cut when position a-1 shares
 * it (cfd[a-1] > L), its last node (of position b-1) when position b does
 * (cfd[b-1] > L).  A cut node gets one partial row per item it spans; a
 * single node cut at either end one.  cnt[k] = the item's partial rows,
 * nf[k] / nl[k] its first / last node, fl[k] = 1 (head cut) | 2 (tail cut). */
__global__ void k_side_count(uint32_t const *__restrict__ start, uint32_t nitems,
                             uint32_t lb, uint32_t L, uint64_t Fb,
                             uint8_t const *__restrict__ cfd, par_ptrs par,
                             uint32_t *__restrict__ cnt, uint32_t *__restrict__ nf,
                             uint32_t *__restrict__ nl, uint32_t *__restrict__ fl)
{
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; k < nitems; k += st) {
    uint32_t const a = start[k], b = start[k + 1];
    if (b <= a) {
      cnt[k] = 0;
      nf[k] = nl[k] = 0;
      fl[k] = 0;
      continue;
    }
    uint32_t x = a, y = b - 1;
    for (uint32_t l = lb; l > L; --l) {
      x = par.p[l][x];
      y = par.p[l][y];
    }
    bool const h = a > 0 && cfd[a - 1] > L;
    bool const t = (uint64_t)b < Fb && cfd[b - 1] > L;
    cnt[k] = (x == y) ? ((h || t) ? 1u : 0u) : (uint32_t)h + (uint32_t)t;
    nf[k] = x;
    nl[k] = y;
    fl[k] = (h ? 1u : 0u) | (t ? 2u : 0u);
  }
}

/* Partial-row slots in item order, the head's before the tail's, from
 * base = exclusive scan of cnt (so the partials of one cut node are
 * consecutive slots); head[k*K+j] / tail[k*K+j] = the slot of item k's
 * head-cut first / tail-cut last node at relative level j (one slot when
 * they are the same node), ~0 when not cut; snode[slot] = its node. */
__global__ void k_side_emit(uint32_t nitems, uint32_t K, uint32_t j,
                            uint32_t const *__restrict__ cnt,
                            uint32_t const *__restrict__ base,
                            uint32_t const *__restrict__ nf,
                            uint32_t const *__restrict__ nl,
                            uint32_t const *__restrict__ fl,
                            uint32_t *__restrict__ head, uint32_t *__restrict__ tail,
                            uint32_t *__restrict__ snode)
{
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; k < nitems; k += st) {
    uint32_t hs = NONE32, ts = NONE32;
    uint32_t const f = fl[k];
    if (cnt[k]) {
      uint32_t q = base[k];
      if (nf[k] == nl[k]) {
        hs = (f & 1u) ? q : NONE32;
        ts = (f & 2u) ? q : NONE32;
        snode[q] = nf[k];
      } else {
        if (f & 1u) {
          hs = q;
          snode[q++] = nf[k];
        }
        if (f & 2u) {
          ts = q;
          snode[q] = nl[k];
        }
      }
    }
    head[k * K + j] = hs;
    tail[k * K + j] = ts;
  }
}

/* runs of equal values: flags[s] = 1 where a run starts */
__global__ void k_run_flags(uint32_t const *__restrict__ v, uint64_t n,
                            uint32_t *__restrict__ flags)
{
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; q < n; q += st) flags[q] = (q == 0 || v[q] != v[q - 1]) ? 1u : 0u;
}

/* run r = its value and first index; first[nrun] = n */
__global__ void k_run_emit(uint32_t const *__restrict__ flags,
                           uint32_t const *__restrict__ pos,
                           uint32_t const *__restrict__ v, uint64_t n, uint32_t nrun,
                           uint32_t *__restrict__ val, uint32_t *__restrict__ first)
{
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; q < n; q += st)
    if (flags[q]) {
      val[pos[q]] = v[q];
      first[pos[q]] = (uint32_t)q;
    }
  if (q == n) first[nrun] = (uint32_t)n; /* the thread just past the end */
}

/* rowptr of ids sorted ascending: first position of every row */
__global__ void k_rowptr(uint32_t const *__restrict__ skey, uint64_t n,
                         uint32_t nrows, uint32_t *__restrict__ rowptr)
{
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; q <= n; q += st) {
    uint32_t const lo = (q == 0) ? 0u : skey[q - 1] + 1u;
    uint32_t const hi = (q == n) ? nrows : skey[q];
    for (uint32_t r = lo; r <= hi && r <= nrows; ++r) rowptr[r] = (uint32_t)q;
  }
}

struct rid_ptrs {
  uint32_t const *par[ADATM_KMAXL];
  uint32_t const *fids[ADATM_KMAXL];
  uint32_t *rid[ADATM_KMAXL];
};

/* rid[l-1][q] = level-(l-1) id of the ancestor of node[q] */
__global__ void k_scatter_rids(uint32_t const *__restrict__ node, uint64_t n,
                               uint32_t d, rid_ptrs rp)
{
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; q < n; q += st) {
    uint32_t v = node[q];
    for (uint32_t l = d; l >= 1; --l) {
      v = rp.par[l][v];
      rp.rid[l - 1][q] = rp.fids[l - 1][v];
    }
  }
}

__global__ void k_compose(uint32_t *__restrict__ cp, uint64_t n1,
                          uint32_t const *__restrict__ fp)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n1; i += st) cp[i] = fp[cp[i]];
}

inline dim3 grid_for(uint64_t n)
{
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return dim3((uint32_t)b);
}

} // namespace

extern "C" {

static int herr(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

uint64_t adatm_k_scan_temp_words(uint64_t n) { return exscan_temp_words(n); }

int adatm_k_exscan(hipStream_t s, uint32_t const *in, uint32_t *out, uint64_t n,
                   uint32_t *temp, uint32_t *total)
{
  hipError_t e = exscan(s, in, out, n, temp);
  if (e == hipSuccess && total) {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, s, in, out, n, total);
    e = hipGetLastError();
  }
  return herr(e);
}

uint64_t adatm_k_sort_temp_words(uint64_t n)
{
  uint64_t const nb = (n + RS_TILE - 1) / RS_TILE;
  uint64_t const hist = 256 * nb;
  return 2 * n + 2 * hist + exscan_temp_words(hist) + 16;
}

/* stable ascending sort of (keys, vals) by the low `nbits` bits of the keys,
 * in place; temp of adatm_k_sort_temp_words(n) words */
int adatm_k_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint64_t n,
                       uint32_t nbits, uint32_t *temp)
{
  if (n <= 1 || nbits == 0) return 0;
  uint32_t const nb = nblk(n, RS_TILE);
  uint64_t const hist = 256ull * nb;
  uint32_t *k2 = temp, *v2 = temp + n;
  uint32_t *gh = temp + 2 * n, *go = gh + hist, *st = go + hist;
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  for (uint32_t shift = 0; shift < nbits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(CB), 0, s, ka, n, shift, nb, gh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return herr(e);
    e = exscan(s, gh, go, hist, st);
    if (e != hipSuccess) return herr(e);
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(CB), 0, s, ka, va, kb, vb,
                       n, shift, nb, go);
    e = hipGetLastError();
    if (e != hipSuccess) return herr(e);
    uint32_t *t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  if (ka != keys) {
    hipError_t e = hipMemcpyAsync(keys, ka, n * 4, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, s);
    return herr(e);
  }
  return 0;
}

int adatm_k_iota(hipStream_t s, uint32_t *p, uint64_t n)
{
  hipLaunchKernelGGL(k_iota, grid_for(n), dim3(256), 0, s, p, n);
  return herr(hipGetLastError());
}

int adatm_k_gather(hipStream_t s, uint32_t *dst, uint32_t const *src,
                   uint32_t const *idx, uint64_t n)
{
  hipLaunchKernelGGL(k_gather, grid_for(n), dim3(256), 0, s, dst, src, idx, n);
  return herr(hipGetLastError());
}

int adatm_k_gather_vals(hipStream_t s, int prec, void *dst, double const *src,
                        uint32_t const *idx, uint64_t n)
{
  if (prec == KP_FP64)
    hipLaunchKernelGGL((k_gather_vals<double>), grid_for(n), dim3(256), 0, s,
                       (double *)dst, src, idx, n);
  else
    hipLaunchKernelGGL((k_gather_vals<float>), grid_for(n), dim3(256), 0, s,
                       (float *)dst, src, idx, n);
  return herr(hipGetLastError());
}

int adatm_k_first_diff(hipStream_t s, uint32_t const *const *sc, uint32_t N,
                       uint64_t n, uint8_t *fd)
{
  sc_ptrs p;
  for (uint32_t k = 0; k < ADATM_KMAXL; ++k) p.p[k] = k < N ? sc[k] : nullptr;
  hipLaunchKernelGGL(k_first_diff, grid_for(n), dim3(256), 0, s, p, N, n, fd);
  return herr(hipGetLastError());
}

int adatm_k_level_flags(hipStream_t s, uint8_t const *fd, uint64_t n, uint32_t l,
                        uint32_t *flags)
{
  hipLaunchKernelGGL(k_level_flags, grid_for(n), dim3(256), 0, s, fd, n, l, flags);
  return herr(hipGetLastError());
}

int adatm_k_level_emit(hipStream_t s, uint32_t const *flags, uint32_t const *pos,
                       uint32_t const *scl, uint64_t n, uint32_t *fids,
                       uint32_t *start)
{
  hipLaunchKernelGGL(k_level_emit, grid_for(n), dim3(256), 0, s, flags, pos, scl,
                     n, fids, start);
  return herr(hipGetLastError());
}

int adatm_k_fptr(hipStream_t s, uint32_t const *start, uint64_t F,
                 uint32_t const *pos_next, uint32_t F_next, uint32_t *fptr)
{
  hipLaunchKernelGGL(k_fptr, grid_for(F + 1), dim3(256), 0, s, start, F, pos_next,
                     F_next, fptr);
  return herr(hipGetLastError());
}

int adatm_k_parent(hipStream_t s, uint32_t const *start_next, uint64_t Fn,
                   uint32_t const *flags, uint32_t const *pos, uint32_t *par)
{
  hipLaunchKernelGGL(k_parent, grid_for(Fn), dim3(256), 0, s, start_next, Fn,
                     flags, pos, par);
  return herr(hipGetLastError());
}

int adatm_k_compose(hipStream_t s, uint32_t *cp, uint64_t n1, uint32_t const *fp)
{
  hipLaunchKernelGGL(k_compose, grid_for(n1), dim3(256), 0, s, cp, n1, fp);
  return herr(hipGetLastError());
}

int adatm_k_items_count(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                        uint32_t tile, uint32_t *nitem, uint32_t *npiece,
                        uint32_t *heavy)
{
  hipLaunchKernelGGL(k_items_count, grid_for(nseg), dim3(256), 0, s, cp, nseg,
                     tile, nitem, npiece, heavy);
  return herr(hipGetLastError());
}

int adatm_k_items_emit(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                       uint32_t tile, uint32_t const *nitem,
                       uint32_t const *ibase, uint32_t const *sbase,
                       uint32_t const *heavy, uint32_t const *hbase,
                       uint32_t *start, uint32_t *seg, uint32_t *slot,
                       uint32_t *hseg, uint32_t *hslot)
{
  hipLaunchKernelGGL(k_items_emit, grid_for(nseg), dim3(256), 0, s, cp, nseg,
                     tile, nitem, ibase, sbase, heavy, hbase, start, seg, slot,
                     hseg, hslot);
  return herr(hipGetLastError());
}

int adatm_k_items_tail(hipStream_t s, uint32_t const *cp, uint64_t nseg,
                       uint32_t nitems, uint32_t nheavy, uint32_t nslots,
                       uint32_t *start, uint32_t *seg, uint32_t *slot,
                       uint32_t *hslot)
{
  hipLaunchKernelGGL(k_items_tail, dim3(1), dim3(1), 0, s, cp, nseg, nitems,
                     nheavy, nslots, start, seg, slot, hslot);
  return herr(hipGetLastError());
}

int adatm_k_fix_count(hipStream_t s, uint32_t const *hslot, uint32_t nheavy,
                      uint32_t ch, uint32_t *nch, uint32_t *multi)
{
  if (nheavy == 0) return 0;
  hipLaunchKernelGGL(k_fix_count, grid_for(nheavy), dim3(256), 0, s, hslot, nheavy,
                     ch, nch, multi);
  return herr(hipGetLastError());
}

int adatm_k_fix_emit(hipStream_t s, uint32_t const *hslot, uint32_t nheavy,
                     uint32_t ch, uint32_t *hchunk, uint32_t nchunk,
                     uint32_t const *multi, uint32_t const *mbase, uint32_t *cstart,
                     uint32_t *cseg, uint32_t *mlist)
{
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, hchunk + nheavy, nchunk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nheavy == 0) return herr(e);
  hipLaunchKernelGGL(k_fix_emit, grid_for(nheavy), dim3(256), 0, s, hslot, nheavy, ch,
                     (uint32_t const *)hchunk, multi, mbase, cstart, cseg, mlist);
  return herr(hipGetLastError());
}

int adatm_k_cfd(hipStream_t s, uint32_t const *const *par, uint32_t lb, uint64_t F,
                uint8_t *cfd)
{
  if (F == 0) return 0;
  if (lb >= ADATM_KMAXL) return (int)hipErrorInvalidValue;
  par_ptrs p;
  for (uint32_t l = 0; l < ADATM_KMAXL; ++l) p.p[l] = (l >= 1 && l <= lb) ? par[l] : nullptr;
  hipLaunchKernelGGL(k_cfd, grid_for(F), dim3(256), 0, s, p, lb, F, cfd);
  return herr(hipGetLastError());
}

int adatm_k_items_anc(hipStream_t s, uint32_t const *start, uint32_t nitems,
                      uint32_t lb, uint32_t lt, uint32_t const *const *par,
                      uint32_t *anc)
{
  par_ptrs p;
  for (uint32_t k = 0; k < ADATM_KMAXL; ++k) p.p[k] = par[k];
  hipLaunchKernelGGL(k_items_anc, grid_for(nitems), dim3(256), 0, s, start, nitems,
                     lb, lt, p, anc);
  return herr(hipGetLastError());
}

int adatm_k_side_count(hipStream_t s, uint32_t const *start, uint32_t nitems,
                       uint32_t lb, uint32_t L, uint64_t Fb, uint8_t const *cfd,
                       uint32_t const *const *par, uint32_t *cnt, uint32_t *nf,
                       uint32_t *nl, uint32_t *fl)
{
  par_ptrs p;
  for (uint32_t k = 0; k < ADATM_KMAXL; ++k) p.p[k] = par[k];
  hipLaunchKernelGGL(k_side_count, grid_for(nitems), dim3(256), 0, s, start, nitems, lb, L,
                     Fb, cfd, p, cnt, nf, nl, fl);
  return herr(hipGetLastError());
}

int adatm_k_side_emit(hipStream_t s, uint32_t nitems, uint32_t K, uint32_t j,
                      uint32_t const *cnt, uint32_t const *base, uint32_t const *nf,
                      uint32_t const *nl, uint32_t const *fl, uint32_t *head,
                      uint32_t *tail, uint32_t *snode)
{
  hipLaunchKernelGGL(k_side_emit, grid_for(nitems), dim3(256), 0, s, nitems, K, j, cnt,
                     base, nf, nl, fl, head, tail, snode);
  return herr(hipGetLastError());
}

int adatm_k_run_flags(hipStream_t s, uint32_t const *v, uint64_t n, uint32_t *flags)
{
  hipLaunchKernelGGL(k_run_flags, grid_for(n), dim3(256), 0, s, v, n, flags);
  return herr(hipGetLastError());
}

int adatm_k_run_emit(hipStream_t s, uint32_t const *flags, uint32_t const *pos,
                     uint32_t const *v, uint64_t n, uint32_t nrun, uint32_t *val,
                     uint32_t *first)
{
  /* one block more than the elements need, so a thread lands on q == n */
  dim3 g = grid_for(n + 1);
  hipLaunchKernelGGL(k_run_emit, g, dim3(256), 0, s, flags, pos, v, n, nrun, val, first);
  return herr(hipGetLastError());
}

int adatm_k_rowptr(hipStream_t s, uint32_t const *skey, uint64_t n, uint32_t nrows,
                   uint32_t *rowptr)
{
  hipLaunchKernelGGL(k_rowptr, grid_for(n + 1), dim3(256), 0, s, skey, n, nrows,
                     rowptr);
  return herr(hipGetLastError());
}

int adatm_k_scatter_rids(hipStream_t s, uint32_t const *node, uint64_t n,
                         uint32_t d, uint32_t const *const *par,
                         uint32_t const *const *fids, uint32_t *const *rid)
{
  rid_ptrs p;
  for (uint32_t k = 0; k < ADATM_KMAXL; ++k) {
    p.par[k] = par[k];
    p.fids[k] = fids[k];
    p.rid[k] = rid[k];
  }
  hipLaunchKernelGGL(k_scatter_rids, grid_for(n), dim3(256), 0, s, node, n, d, p);
  return herr(hipGetLastError());
}

} /* extern "C" */
