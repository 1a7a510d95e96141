"""Relabelling (SPLATT tt_perm / perm_apply / perm_matrix) and dense tiling
(SPLATT_DENSETILE) on the CPU.

SPLATT's reorder.c and tile.c are not part of the reference, so the oracle's
restatement is pinned by hand-worked answers on tiny tensors and by the
property the reference relies on (src/adatm_bench.c:40-51, 529, 654): a
relabelled tensor with its factor rows moved along gives every MTTKRP moved
the same way.  The library (host CSF build, device -1) is then compared
with the oracle on random tensors.
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

MAXM = 88


def _tensor(seed, dims, nnz):
    rng = np.random.default_rng(seed)
    inds = np.stack([rng.integers(0, d, nnz) for d in dims]).astype(np.uint32)
    vals = rng.random(nnz)
    inds, vals, dims2, _ = O.canonicalize(inds, vals, dims)
    return inds, vals, dims2


# ---- the oracle against hand-worked answers ------------------------------
def test_graph_known_answer():
    # vertices: mode 0 -> 0..2, mode 1 -> 3..4, mode 2 -> 5..6
    p = O.perm_graph([3, 2, 2], [2, 0, 1, 1, 0, 0, 0])
    assert [list(x) for x in p] == [[2, 0, 1], [1, 0], [0, 1]]


TINY = np.array([[0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [2, 1, 0, 0, 2]])


def test_fibsched_known_answer():
    # slices of mode 0: i0=0 (part 1), i0=1 (part 0), i0=2 (part 0); the
    # visit reads (1,1,0) (2,0,0) (2,2,2) (0,0,2) (0,2,1)
    p = O.perm_first_touch(TINY, [3, 3, 3], 0, 0, [1, 0, 0])
    assert [list(x) for x in p] == [[2, 0, 1], [1, 0, 2], [0, 2, 1]]


def test_hgraph_known_answer():
    # fibers (i0,i1) in CSF order: (0,0) (0,2) (1,1) (2,0) (2,2) with parts
    # 1 0 1 0 0; the visit reads (0,2,1) (2,0,0) (2,2,2) (0,0,2) (1,1,0)
    p = O.perm_first_touch(TINY, [3, 3, 3], 0, 1, [1, 0, 1, 0, 0])
    assert [list(x) for x in p] == [[0, 2, 1], [1, 2, 0], [1, 0, 2]]


def test_untouched_indices_go_last():
    inds = np.array([[0, 2], [1, 0], [0, 0]])  # index 1 of mode 0, none of 3 of mode 2
    p = O.perm_first_touch(inds, [4, 2, 4], 0, 0, [0, 0])
    assert list(p[0]) == [0, 2, 1, 3]
    assert list(p[2]) == [0, 1, 2, 3]


def test_one_part_in_csf_order_is_identity():
    # a tensor whose CSF order already meets every index in increasing order
    inds = np.array([[0, 0, 1, 1, 2], [0, 1, 1, 2, 2], [0, 1, 2, 2, 3]])
    for level in (0, 1):
        p = O.perm_first_touch(inds, [3, 3, 4], 0, level, [0] * 5)
        assert all(np.array_equal(x, np.arange(len(x))) for x in p)


@pytest.mark.parametrize("kind", ["graph", "hgraph", "fib"])
def test_relabelled_mttkrp_moves_rows(kind):
    """X relabelled, factors shuffled with perm_matrix's rule: every MTTKRP
    is the original one with its rows shuffled the same way."""
    inds, vals, dims = _tensor(3, [30, 20, 12, 7], 3000)
    rng = np.random.default_rng(4)
    mode = 2
    if kind == "graph":
        perms = O.perm_graph(dims, rng.integers(0, 5, sum(dims)))
    else:
        level = len(dims) - 2 if kind == "hgraph" else 0
        F = O.count_fibers(inds, O.basis_order(dims, mode))[level]
        perms = O.perm_first_touch(inds, dims, mode, level, rng.integers(0, 4, F))
    for p, d in zip(perms, dims):
        assert np.array_equal(np.sort(p), np.arange(d))
    U = [rng.random((d, 5)) for d in dims]
    inds2 = np.stack([perms[m][inds[m]] for m in range(len(dims))])
    U2 = [O.shuffle_rows(U[m], perms[m]) for m in range(len(dims))]
    for m in range(len(dims)):
        a = O.coo_mttkrp(inds2, vals, U2, m)
        b = O.shuffle_rows(O.coo_mttkrp(inds, vals, U, m), perms[m])
        assert np.array_equal(a, b)


def test_graph_groups_parts():
    dims = [11, 9, 5]
    parts = np.random.default_rng(8).integers(0, 4, sum(dims))
    perms = O.perm_graph(dims, parts)
    off = 0
    for p, d in zip(perms, dims):
        inv = np.argsort(p)  # new -> old
        assert np.all(np.diff(parts[off + inv]) >= 0)
        off += d


def test_first_touch_labels_count_up_along_the_visit():
    """After relabelling, the visit meets each mode's labels 0, 1, 2, ...
    in order of first appearance."""
    inds, vals, dims = _tensor(5, [25, 18, 9], 800)
    order = O.basis_order(dims, 1)
    F = O.count_fibers(inds, order)
    parts = np.random.default_rng(6).integers(0, 3, F[0])
    perms = O.perm_first_touch(inds, dims, 1, 0, parts)
    # the visit from the CSF of the original tensor
    p = O.lex_order(inds, order)
    s = inds[order][:, p]
    slice_of = np.cumsum(np.r_[True, s[0, 1:] != s[0, :-1]]) - 1
    q = np.argsort(parts[slice_of], kind="stable")
    for l, m in enumerate(order):
        seq = perms[m][s[l, q]]
        firsts = seq[np.sort(np.unique(seq, return_index=True)[1])]
        assert np.array_equal(firsts, np.arange(len(firsts)))


def test_densetile_known_answer():
    inds = np.array([[0, 1, 3, 2], [0, 3, 2, 1], [0, 1, 3, 2]])
    vals = np.array([1.0, 2.0, 3.0, 4.0])
    td, tiles = O.densetile(inds, vals, [4, 4, 4], [0, 1, 2], 2, 1)
    assert td == [1, 2, 2]
    assert [t[3].tolist() for t in tiles] == [[1.0], [4.0], [2.0], [3.0]]
    # the last range takes the rest: index 4 of 5 is in tile 1 (size 2)
    td, tiles = O.densetile(np.array([[0], [4], [4]]), np.ones(1), [5, 5, 5], [0, 1, 2],
                            2, 1)
    assert [t is not None for t in tiles] == [False, False, False, True]


# ---- the library against the oracle -----------------------------------------
@pytest.mark.parametrize("seed,dims,nnz", [(1, [40, 30, 20], 2000),
                                           (2, [25, 20, 15, 10], 4000),
                                           (3, [12, 11, 10, 9, 8], 5000)])
@pytest.mark.parametrize("kind", [A.PERM_GRAPH, A.PERM_HGRAPH, A.PERM_FIBSCHED])
def test_library_labels_match_oracle(seed, dims, nnz, kind):
    inds, vals, dims = _tensor(seed, dims, nnz)
    t = A.Tensor.from_arrays(inds, vals, dims)
    rng = np.random.default_rng(seed + 100)
    for mode in range(len(dims)) if kind != A.PERM_GRAPH else [0]:
        nv = A.perm_nvtxs(t, kind, mode)
        if kind == A.PERM_GRAPH:
            assert nv == sum(dims)
        else:
            F = O.count_fibers(inds, O.basis_order(dims, mode))
            assert nv == F[len(dims) - 2 if kind == A.PERM_HGRAPH else 0]
        parts = rng.integers(0, 1 + mode * 3, nv + 5)  # extra entries are ignored
        got = A.perm_compute(t, kind, parts, mode)
        if kind == A.PERM_GRAPH:
            want = O.perm_graph(dims, parts)
        else:
            level = len(dims) - 2 if kind == A.PERM_HGRAPH else 0
            want = O.perm_first_touch(inds, dims, mode, level, parts)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    perms = A.perm_compute(t, kind, rng.integers(0, 3, A.perm_nvtxs(t, kind, 1)), 1)
    A.relabel(t, perms)
    assert np.array_equal(t.inds(), np.stack([perms[m][inds[m]] for m in range(len(dims))]))


def test_library_rejects_short_partition_and_bad_mode():
    inds, vals, dims = _tensor(9, [10, 8, 6], 200)
    t = A.Tensor.from_arrays(inds, vals, dims)
    with pytest.raises(A.AdatmError, match="partition file has"):
        A.perm_compute(t, A.PERM_GRAPH, np.zeros(sum(dims) - 1))
    with pytest.raises(A.AdatmError, match="basis mode"):
        A.perm_compute(t, A.PERM_HGRAPH, np.zeros(1000), mode=3)
    with pytest.raises(A.AdatmError, match="unknown reordering"):
        A.perm_compute(t, A.PERM_ERROR, np.zeros(1000))


class CsfSparsity(C.Structure):
    _fields_ = [("nfibs", C.c_uint64 * MAXM), ("fptr", C.POINTER(C.c_uint64) * MAXM),
                ("fids", C.POINTER(C.c_uint64) * MAXM), ("vals", C.POINTER(C.c_double))]


class SplattCsf(C.Structure):
    _fields_ = [("nnz", C.c_uint64), ("nmodes", C.c_uint64), ("dims", C.c_uint64 * MAXM),
                ("dim_perm", C.c_uint64 * MAXM), ("which_tile", C.c_int),
                ("ntiles", C.c_uint64), ("tile_dims", C.c_uint64 * MAXM),
                ("pt", C.POINTER(CsfSparsity))]


@pytest.mark.parametrize("alloc,nthreads,depth", [(A.CSF_ONEMODE, 3, 1),
                                                  (A.CSF_TWOMODE, 2, 1),
                                                  (A.CSF_ALLMODE, 4, 2),
                                                  (A.CSF_TWOMODE, 1, 1)])
def test_densetile_csf_matches_oracle(alloc, nthreads, depth):
    """splatt_csf_convert with SPLATT_OPTION_TILE = SPLATT_DENSETILE (host
    build): tile dims, and every tile's fibers, pointers, ids and values."""
    L = A.lib()
    L.splatt_default_opts.restype = C.POINTER(C.c_double)
    L.splatt_free_opts.argtypes = [C.POINTER(C.c_double)]
    L.splatt_csf_convert.argtypes = [C.c_uint64, C.c_uint64,
                                     C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.c_double), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_double)]
    L.splatt_free_csf.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    inds, vals, dims = _tensor(11, [23, 17, 13, 9], 3000)
    N = len(dims)
    opts = L.splatt_default_opts()
    opts[A.OPT_CSF_ALLOC] = alloc
    opts[A.OPT_DEVICE] = -1
    opts[A.OPT_TILE] = A.DENSETILE
    opts[A.OPT_TILEDEPTH] = depth
    opts[A.OPT_NTHREADS] = nthreads
    cols = [np.ascontiguousarray(inds[m], dtype=np.uint64) for m in range(N)]
    ip = (C.POINTER(C.c_uint64) * N)(*[c.ctypes.data_as(C.POINTER(C.c_uint64)) for c in cols])
    out = C.c_void_p()
    assert L.splatt_csf_convert(N, vals.size, ip, vals.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(out), opts) == A.SPLATT_SUCCESS
    ncsf = {A.CSF_ONEMODE: 1, A.CSF_TWOMODE: 2, A.CSF_ALLMODE: N}[alloc]
    arr = C.cast(out, C.POINTER(SplattCsf))
    for g in range(ncsf):
        c = arr[g]
        order = [int(c.dim_perm[l]) for l in range(N)]
        td, tiles = O.densetile(inds, vals, dims, order, nthreads, depth)
        assert c.which_tile == A.DENSETILE
        assert [int(c.tile_dims[m]) for m in range(N)] == td
        assert int(c.ntiles) == len(tiles)
        for k, want in enumerate(tiles):
            pt = c.pt[k]
            F = [int(pt.nfibs[l]) for l in range(N)]
            if want is None:
                assert F == [0] * N
                assert list(np.ctypeslib.as_array(pt.fptr[0], shape=(2,))) == [0, 0]
                continue
            wF, wfptr, wfids, wvals, _ = want
            assert F == wF
            for l in range(N - 1):
                assert np.array_equal(np.ctypeslib.as_array(pt.fptr[l], shape=(F[l] + 1,)),
                                      wfptr[l])
            for l in range(N):
                assert np.array_equal(np.ctypeslib.as_array(pt.fids[l], shape=(F[l],)),
                                      wfids[l])
            assert np.array_equal(np.ctypeslib.as_array(pt.vals, shape=(F[N - 1],)), wvals)
    L.splatt_free_csf(out, opts)
    L.splatt_free_opts(opts)


# ---- the CLI's option handling (no GPU work: every case stops before it) ---
def _cli(*args, cwd):
    return subprocess.run([str(A.CLI_PATH), "bench", *args], capture_output=True,
                          text=True, timeout=60, cwd=cwd)


def test_cli_partition_file_needs_a_known_type(tmp_path):
    tns = tmp_path / "t.tns"
    tns.write_text("1 1 1 1.0\n2 2 2 2.0\n")
    (tmp_path / "p.part").write_text("0\n" * 6)
    r = _cli(str(tns), "-a", "adatm", "-p", "p.part", cwd=tmp_path)
    assert r.returncode != 0
    assert "reordering algorithm '(null)' is not recognized" in r.stderr
    r = _cli(str(tns), "-a", "adatm", "-z", "metis", "-p", "p.part", cwd=tmp_path)
    assert r.returncode != 0
    assert "reordering algorithm 'metis' is not recognized" in r.stderr


def test_cli_bad_partition_file(tmp_path):
    tns = tmp_path / "t.tns"
    tns.write_text("1 1 1 1.0\n2 2 2 2.0\n")
    (tmp_path / "p.part").write_text("0\n1\nx\n")
    r = _cli(str(tns), "-a", "adatm", "-z", "graph", "-p", "p.part", cwd=tmp_path)
    assert r.returncode != 0 and "is not a part id" in r.stderr
    (tmp_path / "p.part").write_text("0\n1\n")
    r = _cli(str(tns), "-a", "adatm", "-z", "graph", "-p", "p.part", cwd=tmp_path)
    assert r.returncode != 0 and "the graph has 6 vertices" in r.stderr
