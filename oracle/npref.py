"""Oracle, numpy half -- TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's algorithms, used by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
Nothing in the product (adatm_amd/) imports this module.

Contents (each function cites the reference it restates):
  coo_mttkrp        brute-force MTTKRP, semantics of mttkrp_stream
                    (src/mttkrp.c:2519-2567): M[i_m] += x * prod_{k!=m} U_k[i_k]
  canonicalize      duplicate merge (sum, draw order) + empty-index compaction
  lex_order / csf   tt_sort + p_mk_outerptr/p_mk_fptr (src/adatm_rcsf.c:21-209)
  count_fibers      count_nfibs (src/adatm_rcsf.c:410-500)
  pair_sort, enumerate_plans, predict, choose
                    src/adatm_sort.c:10-75, src/adatm_bench.c:62-508,
                    src/adatm_rcsf.c:329-406 (planner with its quirks)
  memo_mttkrp_seq   the memoized sequence by its mathematical definition
                    (Z_l bottom-up, Y top-down; src/mttkrp.c:934-1079,
                    1547-1682), computed level by level with numpy
  cpd_als           cpd_als_iterate_adaptive (src/adatm_cpd.c:256-374) with
                    SPLATT's Gram / inverse / normalisation, FP64
  basis_order, perm_graph, perm_first_touch, shuffle_rows
                    SPLATT's tt_perm / perm_apply / perm_matrix (reorder.c of
                    SPLATT v1.1.1, absent from the reference; called at
                    src/cmds/cmd_bench.c:243-251, src/adatm_bench.c:26-51,
                    529, 654), stated by nonzero positions
  densetile         SPLATT_DENSETILE (tile.c / csf.c of SPLATT v1.1.1,
                    absent; selected at src/cmds/cmd_bench.c:151,
                    src/cmds/cmd_cpd.c:90)
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------------------
# brute force
def coo_mttkrp(inds: np.ndarray, vals: np.ndarray, factors, mode: int,
               nrows: int | None = None) -> np.ndarray:
    """M = X_(mode) (khatri-rao of the other factors), FP64, by nonzeros."""
    N = inds.shape[0]
    R = factors[0].shape[1]
    nrows = factors[mode].shape[0] if nrows is None else nrows
    out = np.zeros((nrows, R), dtype=np.float64)
    chunk = 1 << 20
    for s in range(0, vals.shape[0], chunk):
        e = min(s + chunk, vals.shape[0])
        prod = np.repeat(vals[s:e, None].astype(np.float64), R, axis=1)
        for k in range(N):
            if k != mode:
                prod *= factors[k][inds[k, s:e]]
        np.add.at(out, inds[mode, s:e], prod)
    return out


def dense_tensor(inds, vals, dims) -> np.ndarray:
    X = np.zeros(dims, dtype=np.float64)
    np.add.at(X, tuple(inds), vals)
    return X


# ---------------------------------------------------------------------------
# canonical form
def canonicalize(inds: np.ndarray, vals: np.ndarray, dims):
    """Sum duplicates in draw order, drop empty indices, sort by modes 0..N-1.
    Returns (inds, vals, dims, indmaps)."""
    N = inds.shape[0]
    order = np.lexsort(inds[::-1])  # stable; primary key mode 0
    si = inds[:, order]
    sv = vals[order]
    if si.shape[1] == 0:
        return si, sv, list(dims), [np.zeros(0, np.int64)] * N
    new = np.ones(si.shape[1], dtype=bool)
    new[1:] = np.any(si[:, 1:] != si[:, :-1], axis=0)
    starts = np.flatnonzero(new)
    # sequential sum per run (in stable = draw order)
    uv = np.empty(starts.shape[0], dtype=np.float64)
    ends = np.append(starts[1:], si.shape[1])
    # runs of length one are the common case; longer runs summed in order
    uv[:] = sv[starts]
    longr = np.flatnonzero(ends - starts > 1)
    for r in longr:
        acc = 0.0
        for j in range(starts[r], ends[r]):
            acc += sv[j]
        uv[r] = acc
    ui = si[:, starts]
    maps = []
    ndims = []
    for m in range(N):
        occ = np.zeros(dims[m], dtype=bool)
        occ[ui[m]] = True
        newid = np.cumsum(occ) - 1
        ui[m] = newid[ui[m]]
        maps.append(np.flatnonzero(occ))
        ndims.append(int(occ.sum()))
    return ui, uv, ndims, maps


# ---------------------------------------------------------------------------
# CSF
def lex_order(inds: np.ndarray, order) -> np.ndarray:
    """Permutation sorting the coordinates lexicographically by `order`."""
    keys = [inds[m] for m in reversed(list(order))]
    return np.lexsort(keys)


def count_fibers(inds: np.ndarray, order) -> list[int]:
    """F_l = number of distinct length-(l+1) prefixes for l < N-1, and
    F_{N-1} = nnz: every nonzero is a leaf, repeated coordinates included
    (count_nfibs, src/adatm_rcsf.c:410-500, with nfibs[N-1] = nnz set at
    src/adatm_bench.c:181)."""
    N = inds.shape[0]
    p = lex_order(inds, order)
    s = inds[list(order)][:, p]
    F = []
    if s.shape[1] == 0:
        return [0] * N
    diff = np.zeros(s.shape[1], dtype=bool)
    diff[0] = True
    for l in range(N - 1):
        diff[1:] |= s[l, 1:] != s[l, :-1]
        F.append(int(diff.sum()))
    F.append(int(s.shape[1]))
    return F


def csf(inds: np.ndarray, vals: np.ndarray, order):
    """Group CSF: (F, fptr, fids, vals_sorted, par); fids[l][n] = coordinate of
    mode order[l] at node n; fptr[l] = first child; par[l] = parent index."""
    N = inds.shape[0]
    p = lex_order(inds, order)
    s = inds[list(order)][:, p]
    nnz = s.shape[1]
    diff = np.zeros(nnz, dtype=bool)
    diff[0] = True
    starts = []
    for l in range(N):
        diff[1:] |= s[l, 1:] != s[l, :-1]
        starts.append(np.flatnonzero(diff).copy() if l < N - 1 else np.arange(nnz))
    F = [len(x) for x in starts]
    fids = [s[l, starts[l]].astype(np.int64) for l in range(N)]
    fptr, par = [], [None]
    for l in range(N - 1):
        fp = np.searchsorted(starts[l + 1], starts[l]).astype(np.int64)
        fptr.append(np.append(fp, F[l + 1]))
        par.append(np.searchsorted(starts[l], starts[l + 1], side="right") - 1)
    return F, fptr, fids, vals[p].astype(np.float64), par


# ---------------------------------------------------------------------------
# planner (src/adatm_bench.c:62-508, src/adatm_sort.c, src/adatm_utils.c)
def pair_sort(pairs: list[list[int]]) -> list[list[int]]:
    """The reference's in-place quicksort of (x, y) pairs by y, with its exact
    tie behaviour (src/adatm_sort.c:10-67)."""
    a = [list(p) for p in pairs]

    def rng(start, end):
        i, j = start + 1, end - 1
        k = start + (end - start) // 2
        pmid = a[k][:]
        a[k] = a[start][:]
        while i < j:
            if pmid[1] < a[i][1]:
                if pmid[1] > a[j][1]:
                    a[i], a[j] = a[j], a[i]
                    i += 1
                j -= 1
            else:
                if pmid[1] < a[j][1]:
                    j -= 1
                i += 1
        if pmid[1] < a[i][1]:
            i -= 1
        a[start] = a[i][:]
        a[i] = pmid
        if i > start + 1:
            rng(start, i)
        i += 1
        if end - i > 1:
            rng(i, end)

    if len(a) > 1:
        rng(0, len(a))
    return a


def _argmin_pos(arr, con=0):
    mkr = None
    for i, v in enumerate(arr):
        if v > con:
            mkr = i
            break
    for i, v in enumerate(arr):
        if v > con and v < arr[mkr]:
            mkr = i
    return mkr


def _min_pos(arr, con=0):
    vals = [v for v in arr if v > con]
    return min(vals) if vals else None


def product_order(N: int) -> list[list[int]]:
    po = [[0] * (2 * N) for _ in range(N)]
    for m in range(N):
        for i in range(m + 1, N):
            po[m][i] = 1
        for i in range(m):
            po[m][i + N] = 1
    return po


def enumerate_plans(N: int, dims) -> list[list[tuple]]:
    """Candidate plans: list of configs, each a list of groups
    (memo_mode, begin_imten, n_imten, o_opt)."""
    po = product_order(N)
    ngcal = int(math.ceil(N / math.sqrt(2 * (N - 1))))
    configs = []
    for ng in range(ngcal + 1):
        n_grp = ng + 1
        aver = N // n_grp
        rest = N % n_grp
        if aver == 1 and rest > 0:
            continue
        if n_grp >= N:
            continue
        maxim = [0] * (ngcal + 1)
        begin = [0] * (ngcal + 1)
        memo = [0] * (ngcal + 1)
        if n_grp != 1:
            for g in range(n_grp):
                if rest > 0:
                    maxim[g] = aver
                    rest -= 1
                else:
                    maxim[g] = aver - 1
                begin[g] = maxim[g] + 1
            for g in range(1, n_grp):
                memo[g] = memo[g - 1] + maxim[g - 1] + 1
        else:
            maxim[0], maxim[1] = N - 2, 0
            begin[0], begin[1] = N - 1, 1
            memo[0], memo[1] = 0, N - 1
            n_grp = 2
        nvar = _min_pos(maxim, 0)
        for i in range(nvar or 0):
            groups = []
            for g in range(n_grp):
                mm = memo[g]
                b = begin[g]
                nim = maxim[g] - i if mm != N - 1 else 0
                sum_t = [0] * (2 * N)
                for j in range(mm, mm + maxim[g] + 1):
                    for k in range(2 * N):
                        sum_t[k] += po[j][k]
                tmp_t = [sum_t[k] if po[mm][k] > 0 else 0 for k in range(2 * N)]
                tmp_opt = [mm] + [0] * (N - 1)
                for j in range(1, N):
                    k = _argmin_pos(tmp_t, 0)
                    tmp_t[k] = 0
                    tmp_opt[j] = k - N if k >= N else k
                tail = pair_sort([[tmp_opt[j], dims[tmp_opt[j]]] for j in range(b, N)])
                o = [mm] + tmp_opt[1:b] + [x for x, _ in tail]
                groups.append((mm, b, nim, o))
            configs.append(groups)
    return configs


def predict(configs, F_of_order, nnz: int, N: int, R: int, quirks: bool = True):
    """pspace/pops of every config (predict_candidates_adaptive); F_of_order
    maps a level order (tuple) to its fiber counts."""
    V = I = 8
    out = []
    last_F = None
    csf_b = csf_o = 0
    old = None
    for c, groups in enumerate(configs):
        refresh = c == 0 or not quirks or len(groups) != old
        if refresh:
            old = len(groups)
            csf_b = csf_o = 0
        rb = ro = 0
        for g, (mm, b, nim, o) in enumerate(groups):
            if refresh:
                F = F_of_order(tuple(o))
                last_F = F
                csf_b += nnz * V + nnz * I + sum((F[m] + 1) * I + F[m] * I
                                                 for m in range(N - 1))
                csf_o += sum(F[m - 1] * R * 2 for m in range(N, 0, -1))
            else:
                F = last_F
            nre = (groups[g + 1][0] - mm - 1) if g + 1 < len(groups) else N - mm - 1
            for i in range(nim):
                lvl = b - i if quirks else b - 1 - i
                rb += F[lvl] * R * V
            if nim:
                end = b - nim + 1
                for i in range(1, nre + 1):
                    if i < end:
                        ro += F[end - 1] * R * 2
                    elif i < b:
                        ro += F[i] * R * 2
        out.append((csf_b + rb, csf_o + ro))
    return out


def choose(pred, limit=90e9) -> int:
    best = None
    for c, (sp, ops) in enumerate(pred):
        if not sp < limit:
            continue
        if best is None or ops < pred[best][1]:
            best = c
    return best


def plan_use(groups, N):
    grp = [0] * N
    tag = [0] * N
    for g, (mm, _, _, _) in enumerate(groups):
        m1 = groups[g + 1][0] if g + 1 < len(groups) else N
        grp[mm], tag[mm] = g, 0
        for m in range(mm + 1, m1):
            grp[m], tag[m] = g, 1
    return grp, tag


# ---------------------------------------------------------------------------
# memoized sequence by definition
class MemoSeq:
    """The memoized MTTKRP sequence on one plan, FP64, level by level."""

    def __init__(self, inds, vals, dims, groups):
        self.N = inds.shape[0]
        self.dims = list(dims)
        self.groups = groups
        self.csfs = [csf(inds, vals, o) for (_, _, _, o) in groups]
        self.grp, self.tag = plan_use(groups, self.N)
        self.memo = [dict() for _ in groups]

    def _reduce(self, g, l, zin, U):
        """Z_{l-1}(p) = sum_{c in ch(p)} zin(c) o U[fids_l(c)]."""
        F, fptr, fids, xv, _ = self.csfs[g]
        contrib = (zin[:, None] if zin.ndim == 1 else zin) * U[fids[l]]
        seg = fptr[l - 1]
        out = np.add.reduceat(contrib, seg[:-1], axis=0) if contrib.shape[0] else \
            np.zeros((F[l - 1], U.shape[1]))
        return out

    def mttkrp(self, mode, factors):
        g = self.grp[mode]
        mm, b, nim, o = self.groups[g]
        F, fptr, fids, xv, par = self.csfs[g]
        N = self.N
        stored = set(range(b - nim, b)) if nim else set()
        if self.tag[mode] == 0:
            z = xv
            for l in range(N - 1, 0, -1):
                z = self._reduce(g, l, z, factors[o[l]])
                if (l - 1) in stored:
                    self.memo[g][l - 1] = z.copy()
            return z  # level-0 rows == output rows (compacted root)
        d = o.index(mode)
        indeg = min(nim, b - d)
        L = b - indeg
        z = self.memo[g][L]
        for l in range(L, d, -1):
            z = self._reduce(g, l, z, factors[o[l]])
        # Y_{d-1}(parent) top-down
        y = factors[o[0]][np.arange(F[0])]
        for l in range(1, d):
            y = y[par[l]] * factors[o[l]][fids[l]]
        contrib = y[par[d]] * z
        out = np.zeros((self.dims[mode], z.shape[1]))
        np.add.at(out, fids[d], contrib)
        return out


# ---------------------------------------------------------------------------
# CP-ALS (src/adatm_cpd.c:256-374 with SPLATT's dense helpers)
def normalize(A: np.ndarray, first: bool):
    """2-norm columns on the first sweep, else max(1, column max)."""
    if first:
        lam = np.sqrt((A * A).sum(axis=0))
    else:
        lam = np.maximum(A.max(axis=0), 1.0)
    return A / lam, lam


def cpd_als(mttkrp, factors, niters: int, ttnormsq: float, tol: float = 0.0):
    """Generic CP-ALS over an MTTKRP callable mttkrp(mode, factors)."""
    N = len(factors)
    A = [f.copy() for f in factors]
    G = [a.T @ a for a in A]
    fits = []
    lam = np.ones(A[0].shape[1])
    oldfit = 0.0
    for it in range(niters):
        for m in range(N):
            M = mttkrp(m, A)
            V = np.ones_like(G[0])
            for k in range(N):
                if k != m:
                    V *= G[k]
            A[m] = M @ np.linalg.inv(V)
            A[m], lam = normalize(A[m], it == 0)
            G[m] = A[m].T @ A[m]
        H = np.ones_like(G[0])
        for k in range(N):
            H *= G[k]
        norm_mats = abs(lam @ H @ lam)
        inner = ((A[N - 1] * M).sum(axis=0) * lam).sum()
        fit = 1.0 - math.sqrt(ttnormsq + norm_mats - 2.0 * inner) / math.sqrt(ttnormsq)
        fits.append(fit)
        if it > 0 and abs(fit - oldfit) < tol:
            break
        oldfit = fit
    return fits, A, lam


def rel_fro(a: np.ndarray, b: np.ndarray) -> float:
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return float(d / n) if n > 0 else float(d)


# ---------------------------------------------------------------------------
# relabelling (SPLATT reorder.c; callers src/cmds/cmd_bench.c:243-251,
# src/adatm_bench.c:40-51)
def basis_order(dims, mode: int) -> list[int]:
    """Level order of a mode's own CSF: the mode, then the others by
    (dimension, mode)."""
    rest = sorted((m for m in range(len(dims)) if m != mode), key=lambda m: (dims[m], m))
    return [mode] + rest


def perm_graph(dims, parts) -> list[np.ndarray]:
    """Each mode's indices ordered by (part of the index's graph vertex,
    index); vertices are mode 0's indices, then mode 1's, ...  old -> new."""
    out, off = [], 0
    for d in dims:
        key = np.asarray(parts[off:off + d], dtype=np.int64)
        rank = np.empty(d, dtype=np.int64)
        rank[np.argsort(key, kind="stable")] = np.arange(d)
        out.append(rank)
        off += d
    return out


def perm_first_touch(inds, dims, mode: int, level: int, parts) -> list[np.ndarray]:
    """Visit the nonzeros of mode `mode`'s CSF grouped by the part of their
    level-`level` ancestor (stable in CSF order); each mode's new labels go
    to its indices in order of first appearance, unseen indices last in
    index order.  old -> new."""
    N = inds.shape[0]
    order = basis_order(dims, mode)
    p = lex_order(inds, order)
    s = inds[order][:, p]
    nnz = s.shape[1]
    new = np.zeros(nnz, dtype=bool)
    new[0] = True
    for l in range(level + 1):
        new[1:] |= s[l, 1:] != s[l, :-1]
    vertex = np.cumsum(new) - 1
    vparts = np.asarray(parts, dtype=np.int64)[vertex]
    q = np.argsort(vparts, kind="stable")
    out = [None] * N
    for l, m in enumerate(order):
        seq = s[l, q]
        uniq, first = np.unique(seq, return_index=True)
        seen = uniq[np.argsort(first, kind="stable")]
        unseen = np.setdiff1d(np.arange(dims[m]), uniq)
        lab = np.empty(dims[m], dtype=np.int64)
        lab[np.concatenate([seen, unseen]).astype(np.int64)] = np.arange(dims[m])
        out[m] = lab
    return out


def shuffle_rows(A: np.ndarray, perm) -> np.ndarray:
    """Row perm[i] of the result is row i of A (perm_matrix)."""
    B = np.empty_like(A)
    B[np.asarray(perm, dtype=np.int64)] = A
    return B


# ---------------------------------------------------------------------------
# dense tiling (SPLATT_DENSETILE)
def densetile(inds, vals, dims, order, ntile: int, depth: int):
    """Tiles of a CSF with levels at depth >= `depth` cut into `ntile` index
    ranges of size max(dim // ntile, 1) (the last takes the rest).  Returns
    (tile_dims, tiles) with tiles[t] = csf(...) of tile t's nonzeros (or
    None when empty); tile ids are row-major over modes 0..N-1."""
    N = inds.shape[0]
    td = [1] * N
    for d in range(depth, N):
        td[order[d]] = max(ntile, 1)
    tid = np.zeros(inds.shape[1], dtype=np.int64)
    for m in range(N):
        size = max(dims[m] // td[m], 1)
        c = np.minimum(inds[m].astype(np.int64) // size, td[m] - 1)
        tid = tid * td[m] + c
    ntiles = int(np.prod(td))
    tiles = []
    for t in range(ntiles):
        sel = np.flatnonzero(tid == t)
        tiles.append(csf(inds[:, sel], vals[sel], order) if sel.size else None)
    return td, tiles
