"""The SPLATT-compatible C API (include/splatt.h) on the GPU, through ctypes.

splatt_csf_convert / splatt_mttkrp / splatt_cpd_als with every CSF
allocation scheme, compared with the oracle on the same coordinates and
factors (relative Frobenius error <= 1e-12 in FP64), plus the `splatt` CLI.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

pytestmark = pytest.mark.gpu

MAXM = 88


class Kruskal(C.Structure):
    _fields_ = [("rank", C.c_uint64), ("factors", C.POINTER(C.c_double) * MAXM),
                ("lambda_", C.POINTER(C.c_double)), ("nmodes", C.c_uint64),
                ("dims", C.c_uint64 * MAXM), ("fit", C.c_double)]


def _lib():
    L = A.lib()
    L.splatt_default_opts.restype = C.POINTER(C.c_double)
    L.splatt_free_opts.argtypes = [C.POINTER(C.c_double)]
    L.splatt_csf_convert.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.c_double), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_double)]
    L.splatt_mttkrp.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p,
                                C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    L.splatt_cpd_als.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_double),
                                 C.POINTER(Kruskal)]
    L.splatt_free_csf.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.splatt_free_kruskal.argtypes = [C.POINTER(Kruskal)]
    return L


def _convert(L, inds, vals, opts):
    N = inds.shape[0]
    cols = [np.ascontiguousarray(inds[m], dtype=np.uint64) for m in range(N)]
    ip = (C.POINTER(C.c_uint64) * N)(*[c.ctypes.data_as(C.POINTER(C.c_uint64)) for c in cols])
    v = np.ascontiguousarray(vals, dtype=np.float64)
    out = C.c_void_p()
    rc = L.splatt_csf_convert(N, v.shape[0], ip, v.ctypes.data_as(C.POINTER(C.c_double)),
                              C.byref(out), opts)
    assert rc == 1, A.lib().adatm_last_error()
    return out


def _mttkrp(L, csf, mode, U, opts):
    R = U[0].shape[1]
    mats = (C.POINTER(C.c_double) * len(U))(*[u.ctypes.data_as(C.POINTER(C.c_double)) for u in U])
    out = np.zeros((U[mode].shape[0], R))
    rc = L.splatt_mttkrp(mode, R, csf, mats, out.ctypes.data_as(C.POINTER(C.c_double)), opts)
    assert rc == 1, A.lib().adatm_last_error()
    return out


def _tensor(seed, dims, nnz, empty_slices=True, dups=True):
    rng = np.random.default_rng(seed)
    inds = np.stack([rng.integers(0, d, nnz) for d in dims]).astype(np.int64)
    if empty_slices:  # leave holes in every mode
        for m, d in enumerate(dims):
            inds[m] = np.where(inds[m] % 5 == 3, (inds[m] + 1) % d, inds[m])
    if dups:
        inds[:, nnz // 2:nnz // 2 + 50] = inds[:, :50]
    vals = rng.random(nnz)
    return inds, vals


@pytest.mark.parametrize("alloc", [A.ADAPTIVE_TIME_EFFICIENT, A.CSF_ONEMODE,
                                   A.CSF_TWOMODE, A.CSF_ALLMODE])
@pytest.mark.parametrize("dims", [[30, 25, 20], [40, 30, 20, 12], [20, 18, 16, 14, 12],
                                  [9, 8, 9, 8, 9, 8, 9]])
def test_splatt_mttkrp_all_allocations(alloc, dims):
    L = _lib()
    opts = L.splatt_default_opts()
    opts[A.OPT_CSF_ALLOC] = alloc
    inds, vals = _tensor(len(dims) * 7 + alloc, dims, 4000)
    csf = _convert(L, inds, vals, opts)
    R = 12
    rng = np.random.default_rng(1)
    U = [np.ascontiguousarray(rng.random((d, R))) for d in dims]
    for m in range(len(dims)):
        ref = O.coo_mttkrp(inds, vals, U, m)
        err = O.rel_fro(_mttkrp(L, csf, m, U, opts), ref)
        assert err <= 1e-12, (alloc, m, err)
    L.splatt_free_csf(csf, opts)
    L.splatt_free_opts(opts)


def test_memo_revalidated_when_factors_change():
    """Reuse modes called out of sequence, with factors changed in between,
    still match the brute force (the memo is rebuilt when stale)."""
    L = _lib()
    opts = L.splatt_default_opts()
    dims = [25, 20, 15, 10, 8]
    inds, vals = _tensor(3, dims, 6000)
    csf = _convert(L, inds, vals, opts)
    R = 8
    rng = np.random.default_rng(2)
    U = [np.ascontiguousarray(rng.random((d, R))) for d in dims]
    order = [2, 1, 0, 4, 3, 1, 2, 2, 4]
    for k, m in enumerate(order):
        if k % 2 == 1:
            j = int(rng.integers(0, len(dims)))
            U[j][:] = rng.random(U[j].shape)
        err = O.rel_fro(_mttkrp(L, csf, m, U, opts), O.coo_mttkrp(inds, vals, U, m))
        assert err <= 1e-12, (k, m, err)
    L.splatt_free_csf(csf, opts)


@pytest.mark.parametrize("prec,tol", [(A.FP64, 1e-9), (A.FP32, 1e-4)])
def test_splatt_cpd_als(prec, tol):
    L = _lib()
    opts = L.splatt_default_opts()
    opts[A.OPT_NITER] = 8
    opts[A.OPT_TOLERANCE] = 0.0
    opts[A.OPT_RANDSEED] = 99
    opts[A.OPT_VERBOSITY] = 0
    opts[A.OPT_PRECISION] = prec
    dims = [40, 30, 20, 12]
    inds, vals = _tensor(5, dims, 5000)
    csf = _convert(L, inds, vals, opts)
    R = 6
    k = Kruskal()
    assert L.splatt_cpd_als(csf, R, opts, C.byref(k)) == 1
    U0 = [A.rand_matrix(99, m, d, R) for m, d in enumerate(dims)]
    fits, Af, lam = O.cpd_als(lambda m, F: O.coo_mttkrp(inds, vals, F, m), U0, 8,
                              float((vals ** 2).sum()))
    assert abs(k.fit - fits[-1]) <= tol
    # unit columns, lambda absorbs the norms: the Kruskal tensors agree
    got = [np.ctypeslib.as_array(k.factors[m], shape=(dims[m], R)) for m in range(4)]
    for m in range(4):
        np.testing.assert_allclose(np.linalg.norm(got[m], axis=0), 1.0, rtol=1e-6)
    lam_got = np.ctypeslib.as_array(k.lambda_, shape=(R,))
    X1 = np.einsum("r,ar,br,cr,dr->abcd", lam_got, *got)
    X2 = np.einsum("r,ar,br,cr,dr->abcd", lam, *Af)
    assert O.rel_fro(X1, X2) <= (1e-8 if prec == A.FP64 else 1e-3)
    L.splatt_free_kruskal(C.byref(k))
    L.splatt_free_csf(csf, opts)


@pytest.mark.parametrize("dims", [[40, 30, 20, 12], [20, 18, 16, 14, 12], [14, 13, 12, 11, 10, 9]])
def test_space_efficient_allocation(dims):
    """SPACE_EFFICIENT keeps the memoising groups' CSFs only; every mode
    (memo root, reuse, or computed from the leaves) matches the oracle, and
    CP-ALS runs on it."""
    print(f"space efficient {dims}", flush=True)
    L = _lib()
    opts = L.splatt_default_opts()
    opts[A.OPT_CSF_ALLOC] = A.ADAPTIVE_SPACE_EFFICIENT
    inds, vals = _tensor(len(dims) * 3 + 1, dims, 4000)
    csf = _convert(L, inds, vals, opts)
    R = 8
    rng = np.random.default_rng(2)
    U = [np.ascontiguousarray(rng.random((d, R))) for d in dims]
    for m in range(len(dims)):
        ref = O.coo_mttkrp(inds, vals, U, m)
        assert O.rel_fro(_mttkrp(L, csf, m, U, opts), ref) <= 1e-12, m
    k = Kruskal()
    opts[A.OPT_NITER] = 3
    opts[A.OPT_TOLERANCE] = 0.0
    opts[A.OPT_VERBOSITY] = 0
    assert L.splatt_cpd_als(csf, R, opts, C.byref(k)) == 1
    assert 0.0 < k.fit < 1.0
    L.splatt_free_kruskal(C.byref(k))
    L.splatt_free_csf(csf, opts)
    L.splatt_free_opts(opts)


@pytest.mark.parametrize("planner", [A.PLAN_LEARNED, A.PLAN_EXHAUSTIVE])
def test_planner_options(planner, tmp_path, monkeypatch):
    """The learned (analytic fallback without a model) and exhaustive planners
    behind splatt_csf_convert give CSF sets whose MTTKRPs match the oracle."""
    print(f"planner option {planner}", flush=True)
    monkeypatch.setenv("ADATM_COSTMODEL", str(tmp_path / "missing.bin"))
    L = _lib()
    inds, vals = _tensor(5, [40, 30, 20, 12], 6000)
    opts = L.splatt_default_opts()
    opts[A.OPT_PLANNER] = planner
    opts[A.OPT_RANK_HINT] = 8
    csf = _convert(L, inds, vals, opts)
    rng = np.random.default_rng(1)
    U = [np.ascontiguousarray(rng.random((d, 8))) for d in [40, 30, 20, 12]]
    for m in range(4):
        ref = O.coo_mttkrp(inds, vals, U, m)
        assert O.rel_fro(_mttkrp(L, csf, m, U, opts), ref) <= 1e-12
    L.splatt_free_csf(csf, opts)
    L.splatt_free_opts(opts)


def test_cli_bench_and_cpd(tmp_path):
    exe = str(A.CLI_PATH)
    print("cli: bench on the full 4-mode workload", flush=True)
    r = subprocess.run([exe, "bench", "synth:1", "-a", "adatm", "-i", "2", "-r", "16"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "CSF-STORAGE" in r.stdout and "vCSF-STORAGE" in r.stdout
    assert "NNZ=2000000" in r.stdout
    assert r.stdout.count("  mode ") == 8 and "its =   2" in r.stdout
    r = subprocess.run([exe, "cpd", "synth:1:200000", "-r", "8", "-i", "5", "--tol", "0",
                        "--seed", "3"], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Final fit:" in r.stdout and r.stdout.count("  its = ") == 5
    assert (tmp_path / "lambda.mat").exists() and (tmp_path / "mode4.mat").exists()
    M = np.loadtxt(tmp_path / "mode1.mat")
    assert M.shape[1] == 8
    np.testing.assert_allclose(np.linalg.norm(M, axis=0), 1.0, rtol=1e-6)
    for planner, label in (("learned", "Planner: "), ("exhaustive", "Planner: exhaustive")):
        r = subprocess.run([exe, "bench", "synth:1:100000", "-a", "adatm", "-i", "1", "-r", "8",
                            "--precision", "fp32", "--planner", planner],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert label in r.stdout and "its =   1" in r.stdout


def test_cli_outputs_match_the_oracle(tmp_path):
    """`splatt bench -a adatm -a csf -w` writes each algorithm's MTTKRPs
    (factors from seed 1) and `splatt cpd` its fits and factors: both against
    the oracle on the canonical form of the fixture file (the CLI sums
    duplicates and drops empty indices on input).  The reference's
    reordering options need a partition file, -s scales the thread count
    by 2."""
    exe = str(A.CLI_PATH)
    tns = os.path.join(os.path.dirname(__file__), "golden", "small4.tns")
    a = np.loadtxt(tns, comments="#", ndmin=2)
    inds0 = a[:, :4].T.astype(np.int64) - 1
    ci, cv, cd, _ = O.canonicalize(inds0, a[:, 4].copy(), [int(x) + 1 for x in inds0.max(axis=1)])
    R = 8
    r = subprocess.run([exe, "bench", tns, "-a", "adatm", "-a", "csf", "-w", "-i", "1",
                        "-r", str(R)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    U = [A.rand_matrix(1, m, d, R) for m, d in enumerate(cd)]
    for alg in ("adatm", "csf"):
        for m in range(4):
            M = np.loadtxt(tmp_path / f"{alg}_mode{m + 1}.mat").reshape(cd[m], R)
            assert O.rel_fro(M, O.coo_mttkrp(ci, cv, U, m)) <= 1e-7, (alg, m)
    r = subprocess.run([exe, "cpd", tns, "-r", str(R), "-i", "10", "--tol", "0", "--seed", "3"],
                       capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    fits = [float(l.split("fit = ")[1].split()[0]) for l in r.stdout.splitlines()
            if "fit = " in l]
    U3 = [A.rand_matrix(3, m, d, R) for m, d in enumerate(cd)]
    ofits, oA, olam = O.cpd_als(lambda m, F: O.coo_mttkrp(ci, cv, F, m), U3, 10,
                                float((cv ** 2).sum()))
    assert len(fits) == 10 and np.max(np.abs(np.array(fits) - ofits)) <= 1e-5
    lam = np.loadtxt(tmp_path / "lambda.mat")
    for m in range(4):
        s_ = np.sqrt((oA[m] ** 2).sum(axis=0))
        olam = olam * s_
        F = np.loadtxt(tmp_path / f"mode{m + 1}.mat")
        assert O.rel_fro(F, oA[m] / s_) <= 1e-6, m
    assert O.rel_fro(lam, olam) <= 1e-6
    r = subprocess.run([exe, "bench", tns, "-a", "adatm", "-p", "part.txt"],
                       capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode != 0 and "algorithm '(null)' is not recognized" in r.stderr
    for opt in (["-z", "graph"], ["-m", "2"]):  # no partition file: no reordering
        r = subprocess.run([exe, "bench", tns, "-a", "adatm", "-i", "1", *opt],
                           capture_output=True, text=True, timeout=60, cwd=tmp_path)
        assert r.returncode == 0 and "Reordering" not in r.stdout, (opt, r.stderr)
    r = subprocess.run([exe, "bench", tns, "-a", "splatt"], capture_output=True, text=True,
                       timeout=60, cwd=tmp_path)
    assert r.returncode != 0 and "not recognized" in r.stderr
    r = subprocess.run([exe, "bench", tns, "-a", "adatm", "-s", "-t", "4", "-i", "1",
                        "--tile"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert [l for l in r.stdout.splitlines() if l.startswith("## THREADS")] == [
        "## THREADS 1", "## THREADS 2", "## THREADS 4"]


class CsfSparsity(C.Structure):
    _fields_ = [("nfibs", C.c_uint64 * MAXM), ("fptr", C.POINTER(C.c_uint64) * MAXM),
                ("fids", C.POINTER(C.c_uint64) * MAXM), ("vals", C.POINTER(C.c_double))]


class SplattCsf(C.Structure):
    _fields_ = [("nnz", C.c_uint64), ("nmodes", C.c_uint64), ("dims", C.c_uint64 * MAXM),
                ("dim_perm", C.c_uint64 * MAXM), ("which_tile", C.c_int),
                ("ntiles", C.c_uint64), ("tile_dims", C.c_uint64 * MAXM),
                ("pt", C.POINTER(CsfSparsity))]


def _csf_arrays(ptr, n):
    """(dim_perm, nfibs, fptr, fids, vals) of the n CSFs of a SPLATT array"""
    arr = C.cast(ptr, C.POINTER(SplattCsf))
    out = []
    for g in range(n):
        c = arr[g]
        N, nnz = int(c.nmodes), int(c.nnz)
        pt = c.pt[0]
        F = [int(pt.nfibs[l]) for l in range(N)]
        fptr = [np.ctypeslib.as_array(pt.fptr[l], shape=(F[l] + 1,)).copy()
                for l in range(N - 1)]
        fids = [np.ctypeslib.as_array(pt.fids[l], shape=(F[l],)).copy() for l in range(N)]
        vals = np.ctypeslib.as_array(pt.vals, shape=(nnz,)).copy()
        out.append(([int(c.dim_perm[l]) for l in range(N)], F, fptr, fids, vals))
    return out


@pytest.mark.parametrize("alloc", [A.ADAPTIVE_TIME_EFFICIENT, A.CSF_TWOMODE, A.CSF_ALLMODE])
def test_device_built_csf_equals_host_built(alloc):
    """splatt_csf_convert builds the SPLATT arrays on the GPU (radix sort and
    scans) by default: plan, fptr, fids and value order are bit-identical to
    the host build's (device -1), with duplicates and empty slices."""
    print(f"device vs host CSF alloc={alloc}", flush=True)
    L = _lib()
    dims = [60, 45, 30, 20, 9]
    inds, vals = _tensor(17, dims, 50000)
    built = []
    for dev in (0, -1):
        opts = L.splatt_default_opts()
        opts[A.OPT_CSF_ALLOC] = alloc
        opts[A.OPT_DEVICE] = dev
        csf = _convert(L, inds, vals, opts)
        n = {A.CSF_TWOMODE: 2, A.CSF_ALLMODE: 5}.get(alloc)
        if n is None:
            t = A.Tensor.from_arrays(inds, vals, [int(x) + 1 for x in inds.max(axis=1)])
            cands, ch = A.plan_analytic(t, 16, device=-1)
            n = cands[ch].n_grp
        built.append(_csf_arrays(csf, n))
        L.splatt_free_csf(csf, opts)
        L.splatt_free_opts(opts)
    assert len(built[0]) == len(built[1])
    for a, b in zip(*built):
        assert a[0] == b[0] and a[1] == b[1]
        for x, y in zip(a[2] + a[3] + [a[4]], b[2] + b[3] + [b[4]]):
            assert np.array_equal(x, y)


def _canonical_fixture():
    tns = os.path.join(os.path.dirname(__file__), "golden", "small4.tns")
    a = np.loadtxt(tns, comments="#", ndmin=2)
    inds0 = a[:, :4].T.astype(np.int64) - 1
    ci, cv, cd, _ = O.canonicalize(inds0, a[:, 4].copy(),
                                   [int(x) + 1 for x in inds0.max(axis=1)])
    return tns, ci, cv, cd


@pytest.mark.parametrize("kind,mode", [("graph", 1), ("hgraph", 2), ("fib", 1), ("fib", 4)])
def test_cli_reordering_keeps_outputs(tmp_path, kind, mode):
    """`splatt bench -z KIND -p FILE -m MODE -w`: the tensor is relabelled,
    the factors follow their rows, and the MTTKRPs written back in the
    original labels equal the oracle's on the unrelabelled tensor; -a csf
    with --tile (dense tiling, 3 tiles per tiled level) as well."""
    print(f"cli reorder kind={kind} mode={mode}", flush=True)
    exe = str(A.CLI_PATH)
    tns, ci, cv, cd = _canonical_fixture()
    if kind == "graph":
        nv = sum(cd)
    else:
        F = O.count_fibers(ci, O.basis_order(cd, mode - 1))
        nv = F[len(cd) - 2] if kind == "hgraph" else F[0]
    parts = np.random.default_rng(mode).integers(0, 7, nv)
    np.savetxt(tmp_path / "p.part", parts, fmt="%d")
    R = 8
    r = subprocess.run([exe, "bench", tns, "-a", "adatm", "-a", "csf", "-w", "-i", "1",
                        "-r", str(R), "-z", kind, "-p", "p.part", "-m", str(mode), "-t", "3",
                        "--tile"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert f"TYPE={kind}" in r.stdout
    U = [A.rand_matrix(1, m, d, R) for m, d in enumerate(cd)]
    for alg in ("adatm", "csf"):
        for m in range(4):
            M = np.loadtxt(tmp_path / f"{alg}_mode{m + 1}.mat").reshape(cd[m], R)
            assert O.rel_fro(M, O.coo_mttkrp(ci, cv, U, m)) <= 1e-7, (alg, m)


@pytest.mark.parametrize("kind", [A.PERM_HGRAPH, A.PERM_FIBSCHED])
def test_device_relabel_matches_oracle(kind):
    """The relabelling with the basis CSF built on the GPU equals the oracle,
    on a tensor with duplicates and empty indices."""
    print(f"device relabel kind={kind}", flush=True)
    dims = [50, 40, 30, 20, 10]
    inds, vals = _tensor(21, dims, 40000)
    t = A.Tensor.from_arrays(inds, vals, dims)
    rng = np.random.default_rng(kind)
    for mode in range(len(dims)):
        nv = A.perm_nvtxs(t, kind, mode, device=0)
        parts = rng.integers(0, 16, nv)
        got = A.perm_compute(t, kind, parts, mode, device=0)
        level = len(dims) - 2 if kind == A.PERM_HGRAPH else 0
        want = O.perm_first_touch(inds.astype(np.int64), dims, mode, level, parts)
        for g, w in zip(got, want):
            assert np.array_equal(g, w), mode


class Sptensor(C.Structure):
    _fields_ = [("nnz", C.c_uint64), ("nmodes", C.c_uint64), ("dims", C.POINTER(C.c_uint64)),
                ("ind", C.POINTER(C.POINTER(C.c_uint64))), ("vals", C.POINTER(C.c_double)),
                ("tiled", C.c_int), ("indmap", C.POINTER(C.c_uint64) * MAXM)]


class Permutation(C.Structure):
    _fields_ = [("perms", C.POINTER(C.c_uint64) * MAXM),
                ("iperms", C.POINTER(C.c_uint64) * MAXM)]


def test_tt_perm_relabels_in_place(tmp_path):
    """SPLATT's tt_perm on a tensor from tt_read (duplicates kept): perms and
    iperms are inverse, the indices are relabelled, and perm_matrix moves
    row i to row perms[i]."""
    L = _lib()
    L.tt_read.restype = C.POINTER(Sptensor)
    L.tt_read.argtypes = [C.c_char_p]
    L.tt_free.argtypes = [C.POINTER(Sptensor)]
    L.tt_perm.restype = C.POINTER(Permutation)
    L.tt_perm.argtypes = [C.POINTER(Sptensor), C.c_int, C.c_uint64, C.c_char_p]
    L.perm_free.argtypes = [C.POINTER(Permutation)]
    tns, _, _, _ = _canonical_fixture()
    a = np.loadtxt(tns, comments="#", ndmin=2)
    inds = a[:, :4].T.astype(np.int64) - 1
    dims = [int(x) + 1 for x in inds.max(axis=1)]
    mode = 2
    F = O.count_fibers(inds, O.basis_order(dims, mode))
    parts = np.random.default_rng(2).integers(0, 5, F[2])
    np.savetxt(tmp_path / "h.part", parts, fmt="%d")
    tt = L.tt_read(tns.encode())
    assert tt
    p = L.tt_perm(tt, A.PERM_HGRAPH, mode, str(tmp_path / "h.part").encode())
    assert p, A.lib().adatm_last_error()
    want = O.perm_first_touch(inds, dims, mode, 2, parts)
    for m in range(4):
        perm = np.ctypeslib.as_array(p.contents.perms[m], shape=(dims[m],))
        iperm = np.ctypeslib.as_array(p.contents.iperms[m], shape=(dims[m],))
        assert np.array_equal(perm, want[m])
        assert np.array_equal(iperm[perm], np.arange(dims[m]))
        got = np.ctypeslib.as_array(tt.contents.ind[m], shape=(inds.shape[1],))
        assert np.array_equal(got, want[m][inds[m]])
    L.perm_free(p)
    L.tt_free(tt)


@pytest.mark.parametrize("alloc", [A.CSF_ONEMODE, A.CSF_ALLMODE])
def test_densetile_device_equals_host_and_mttkrp(alloc):
    """A DENSETILE CSF set built on the GPU equals the host build tile by
    tile, and splatt_mttkrp on it equals the oracle (FP64, 1e-12)."""
    print(f"densetile alloc={alloc}", flush=True)
    L = _lib()
    dims = [40, 30, 20, 12]
    inds, vals = _tensor(31, dims, 20000)
    N = len(dims)
    built = []
    for dev in (0, -1):
        opts = L.splatt_default_opts()
        opts[A.OPT_CSF_ALLOC] = alloc
        opts[A.OPT_DEVICE] = dev
        opts[A.OPT_TILE] = A.DENSETILE
        opts[A.OPT_NTHREADS] = 4
        csf = _convert(L, inds, vals, opts)
        arr = C.cast(csf, C.POINTER(SplattCsf))
        tiles = []
        for g in range({A.CSF_ONEMODE: 1, A.CSF_ALLMODE: N}[alloc]):
            c = arr[g]
            assert c.which_tile == A.DENSETILE and c.ntiles == 4 ** (N - 1)
            for k in range(int(c.ntiles)):
                pt = c.pt[k]
                F = [int(pt.nfibs[l]) for l in range(N)]
                if F[0] == 0:
                    tiles.append(None)
                    continue
                tiles.append((F, [np.ctypeslib.as_array(pt.fptr[l], shape=(F[l] + 1,)).copy()
                                  for l in range(N - 1)],
                              [np.ctypeslib.as_array(pt.fids[l], shape=(F[l],)).copy()
                               for l in range(N)],
                              np.ctypeslib.as_array(pt.vals, shape=(F[N - 1],)).copy()))
        built.append(tiles)
        if dev == 0:
            rng = np.random.default_rng(5)
            U = [rng.random((d, 8)) for d in dims]
            for m in range(N):
                got = _mttkrp(L, csf, m, U, opts)
                assert O.rel_fro(got, O.coo_mttkrp(inds.astype(np.int64), vals, U, m)) <= 1e-12
        L.splatt_free_csf(csf, opts)
        L.splatt_free_opts(opts)
    assert len(built[0]) == len(built[1])
    for a, b in zip(*built):
        assert (a is None) == (b is None)
        if a is None:
            continue
        assert a[0] == b[0]
        for x, y in zip(a[1] + a[2] + [a[3]], b[1] + b[2] + [b[3]]):
            assert np.array_equal(x, y)
