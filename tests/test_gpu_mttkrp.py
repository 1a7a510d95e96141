"""GPU parity of the memoized MTTKRP sequence against the oracle.

Every mode's MTTKRP from the HIP path (through the C ABI) is compared with
the brute-force COO MTTKRP of the oracle (mttkrp_stream semantics) on the
same canonical tensor and factors.  Tolerances: relative Frobenius error
<= 1e-12 for FP64, <= 1e-5 for FP32, <= 1e-2 for BF16 factor rows.
"""
import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

pytestmark = pytest.mark.gpu

TOL = {A.FP64: 1e-12, A.FP32: 1e-5, A.BF16ROWS: 2e-2}


def _tensor(dims, ndraws, dist, seed, values=A.VALS_U01):
    spec = A.make_spec(dims, ndraws, dist=dist, zipf_s=1.1, nclusters=50, sigma=3.0,
                       values=values, seed=seed)
    t = A.Tensor.synth(spec)
    t.canonicalize()
    return t


def _factors(t, R, seed=11):
    return [A.rand_matrix(seed, m, d, R) for m, d in enumerate(t.dims)]


CASES = [
    # dims, ndraws, dist, R
    ([40, 30, 20], 3000, A.DIST_UNIFORM, 8),
    ([50, 40, 30, 20], 20000, A.DIST_UNIFORM, 16),
    ([200, 150, 100, 50, 30], 50000, A.DIST_ZIPF, 32),
    ([30, 30, 30, 30, 30, 30], 20000, A.DIST_CLUSTER, 10),
    ([20] * 7, 15000, A.DIST_UNIFORM, 16),
    ([16] * 8, 30000, A.DIST_CLUSTER, 64),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec", [A.FP64, A.FP32])
@pytest.mark.parametrize("staged", [False, True])
def test_every_plan_every_mode(case, prec, staged):
    """Every candidate plan and mode against the COO brute force, with the
    memo levels written beside one leaf-to-root pass (default; small tiles so
    that many memo nodes are cut by work items) or by a chain of passes."""
    dims, nd, dist, R = CASES[case]
    print(f"case={case} dims={dims} R={R} prec={prec} staged={staged}", flush=True)
    t = _tensor(dims, nd, dist, seed=100 + case)
    U = _factors(t, R)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(t.nmodes)]
    for plan in A.plan_candidates(t.nmodes, t.dims):
        e = A.Engine(t, R, plan, precision=prec, staged=staged,
                     tiles=(0, 0) if staged else (16, 0))
        for m in range(t.nmodes):
            e.set_factor(m, U[m])
        for m in range(t.nmodes):
            M = e.mttkrp(m)
            err = O.rel_fro(M, ref[m])
            assert err <= TOL[prec], (plan.groups(t.nmodes), m, err)
        e.close()


@pytest.mark.parametrize("N,prec,R", [(10, A.FP64, 16), (10, A.FP32, 32), (12, A.FP64, 16),
                                      (16, A.FP64, 8), (16, A.FP32, 32)])
def test_many_modes_every_plan(N, prec, R):
    """Tensors of 10, 12 and 16 modes (the engine's limit): leaf-to-root
    passes are cut at 8 fused levels and reuse scatters reach 8, 10 and 14
    levels deep (levels 9..16 share one kernel instance whose depth is a
    runtime bound); every candidate plan and mode against the COO brute
    force."""
    print(f"many modes N={N} prec={prec} R={R}", flush=True)
    dims = [6 + (m % 3) for m in range(N)]
    t = _tensor(dims, 20000, A.DIST_CLUSTER, seed=200 + N)
    U = _factors(t, R)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(N)]
    for plan in A.plan_candidates(N, t.dims):
        e = A.Engine(t, R, plan, precision=prec)
        for m in range(N):
            e.set_factor(m, U[m])
        for m in range(N):
            err = O.rel_fro(e.mttkrp(m), ref[m])
            assert err <= TOL[prec], (plan.groups(N), m, err)
        e.close()


def test_bf16_rows():
    dims, nd, dist, R = CASES[2]
    t = _tensor(dims, nd, dist, seed=7)
    U = _factors(t, R)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    plan = A.plan_candidates(t.nmodes, t.dims)[0]
    e = A.Engine(t, R, plan, precision=A.BF16ROWS)
    for m in range(t.nmodes):
        e.set_factor(m, U[m])
    for m in range(t.nmodes):
        err = O.rel_fro(e.mttkrp(m), O.coo_mttkrp(inds, vals, U, m))
        assert err <= TOL[A.BF16ROWS], (m, err)


def test_heavy_segments_split():
    """A few very long fibers (more children than one work item holds) are
    cut into pieces and summed by the fix-up pass."""
    rng = np.random.default_rng(3)
    nnz = 40000
    inds = np.stack([np.zeros(nnz, np.int64), rng.integers(0, 2, nnz),
                     rng.integers(0, 4, nnz), rng.integers(0, 5000, nnz)])
    inds[0, :100] = rng.integers(0, 30, 100)
    vals = rng.random(nnz)
    t = A.Tensor.from_arrays(inds, vals, [30, 2, 4, 5000])
    t.canonicalize()
    R = 16
    U = _factors(t, R)
    ii, vv = t.inds().astype(np.int64), t.vals.copy()
    for plan in A.plan_candidates(4, t.dims):
        e = A.Engine(t, R, plan, precision=A.FP64)
        for m in range(4):
            e.set_factor(m, U[m])
        for m in range(4):
            err = O.rel_fro(e.mttkrp(m), O.coo_mttkrp(ii, vv, U, m))
            assert err <= 1e-12, (plan.groups(4), m, err)


def test_memo_is_overwritten_between_sweeps():
    """Repeated sequences with changing factors stay exact (the reference
    accumulates into its memo, mttkrp.c:967/1022, and drifts)."""
    dims, nd, dist, R = CASES[1]
    t = _tensor(dims, nd, dist, seed=5)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    plan = A.plan_candidates(4, t.dims)[0]
    e = A.Engine(t, R, plan, precision=A.FP64)
    for rep in range(3):
        U = _factors(t, R, seed=40 + rep)
        for m in range(4):
            e.set_factor(m, U[m])
        for m in range(4):
            err = O.rel_fro(e.mttkrp(m), O.coo_mttkrp(inds, vals, U, m))
            assert err <= 1e-12, (rep, m, err)


def test_csf_structure_bit_exact():
    dims, nd, dist, R = CASES[2]
    t = _tensor(dims, nd, dist, seed=9)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    plan = A.plan_candidates(t.nmodes, t.dims)[-1]
    e = A.Engine(t, R, plan, precision=A.FP64)
    for g, grp in enumerate(plan.groups(t.nmodes)):
        F, fptr, fids, gv = e.get_csf(g, t.nnz)
        Fo, fpo, fio, vo, _ = O.csf(inds, vals, grp[3])
        assert F == Fo
        for l in range(t.nmodes - 1):
            assert np.array_equal(fptr[l], fpo[l])
        for l in range(1, t.nmodes):
            assert np.array_equal(fids[l], fio[l])
        assert np.array_equal(gv, vo)


@pytest.mark.parametrize("prec", [A.FP64, A.FP32])
@pytest.mark.parametrize("graph", [False, True])
def test_cpd_fit_matches_oracle(prec, graph):
    dims, nd, dist, R = CASES[1]
    t = _tensor(dims, nd, dist, seed=21)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    U = _factors(t, R, seed=77)
    cands, ch = A.plan_analytic(t, R)
    e = A.Engine(t, R, cands[ch], precision=prec)
    for m in range(4):
        e.set_factor(m, U[m])
    fits = e.cpd(10, tol=0.0, graph=graph)
    ref, _, _ = O.cpd_als(lambda m, F: O.coo_mttkrp(inds, vals, F, m), U, 10,
                          float((vals ** 2).sum()))
    tol = 1e-9 if prec == A.FP64 else 1e-4
    assert np.max(np.abs(np.asarray(fits) - np.asarray(ref))) <= tol, (fits, ref)


@pytest.mark.parametrize("R", [10, 40, 64])
def test_cpd_fit_ranks_fp32(R):
    """FP32 sweeps at ranks that pad one (R=10) or use two (R=40, 64) 32-wide
    Gram tiles on the matrix cores match the FP64 oracle's fits."""
    print(f"cpd ranks R={R}", flush=True)
    t = _tensor([90, 80, 70, 66], 40000, A.DIST_ZIPF, seed=23)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    U = _factors(t, R, seed=78)
    cands, ch = A.plan_analytic(t, R)
    e = A.Engine(t, R, cands[ch], precision=A.FP32)
    for m in range(4):
        e.set_factor(m, U[m])
    fits = e.cpd(6, tol=0.0, graph=True)
    ref, _, _ = O.cpd_als(lambda m, F: O.coo_mttkrp(inds, vals, F, m), U, 6,
                          float((vals ** 2).sum()))
    assert np.max(np.abs(np.asarray(fits) - np.asarray(ref))) <= 1e-4, (fits, ref)


@pytest.mark.parametrize("R", [10, 20, 40, 64])
def test_cpd_fit_ranks_fp64(R):
    """FP64 sweeps whose Grams run on the FP64 matrix cores in 1, 2, 3 and 4
    column tiles of 16 (R = 10, 20, 40, 64) match the oracle's fits to 1e-9."""
    print(f"cpd fp64 ranks R={R}", flush=True)
    t = _tensor([90, 80, 70, 66], 40000, A.DIST_ZIPF, seed=24)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    U = _factors(t, R, seed=79)
    cands, ch = A.plan_analytic(t, R)
    for graph in (False, True):
        e = A.Engine(t, R, cands[ch], precision=A.FP64)
        for m in range(4):
            e.set_factor(m, U[m])
        fits = e.cpd(6, tol=0.0, graph=graph)
        ref, _, _ = O.cpd_als(lambda m, F: O.coo_mttkrp(inds, vals, F, m), U, 6,
                              float((vals ** 2).sum()))
        assert np.max(np.abs(np.asarray(fits) - np.asarray(ref))) <= 1e-9, (graph, fits, ref)
        e.close()


def test_tiles_and_autotuner():
    """Work-item tiles change only the summation grouping: every tile gives
    the oracle's MTTKRP; the autotuner's choice is one of its candidates."""
    print("tiles and autotuner", flush=True)
    t = _tensor([300, 200, 100, 50], 60000, A.DIST_ZIPF, seed=29)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    R = 16
    U = _factors(t, R, seed=5)
    cands, ch = A.plan_analytic(t, R)
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(4)]
    for tiles in ((8, 8, 4), (64, 32, 16), (512, 256, 8), (4096, 4096, 0)):
        e = A.Engine(t, R, cands[ch], precision=A.FP64, tiles=tiles)
        for m in range(4):
            e.set_factor(m, U[m])
        for m in range(4):
            assert O.rel_fro(e.mttkrp(m), ref[m]) <= 1e-12, (tiles, m)
        e.close()
    (rt, st, bt, ft), ms = A.autotune_tiles(t, R, cands[ch], A.FP32, 0, 2)
    assert rt in (32, 64, 128, 256, 512) and st in (16, 32, 64, 128, 256)
    assert bt & ~(A.BATCH_PIPELINED | A.BATCH_WIDE) in (0, 4, 8, 16) and ms > 0
    groups = sum(1 << g for g in range(cands[ch].n_grp) if cands[ch].grp[g].n_imten)
    assert ft & ~sum(A.ENGINE_STAGED_GROUP(g) for g in range(16) if groups >> g & 1) == 0


@pytest.mark.parametrize("prec,R", [(A.FP64, 16), (A.FP32, 16), (A.FP32, 32), (A.FP32, 48)])
def test_every_batch_every_plan(prec, R):
    """Children per reduction batch (4/8/16, plain or pipelined) change only
    when rows are loaded, never the summation order: every batch value gives
    the oracle's MTTKRP on every plan, and all agree bit for bit.  R = 48
    runs the one-group-per-wave kernels (lane values read as scalars); the
    wide batches put 16-byte row pieces in each lane (FP64 R = 16: 8 lanes of
    2 columns; FP32 R = 32 / 48: 8 / 12 lanes of 4 columns)."""
    print(f"batches prec={prec} R={R}", flush=True)
    t = _tensor([120, 90, 60, 40, 20, 10], 30000, A.DIST_ZIPF, seed=31)
    U = _factors(t, R, seed=9)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(t.nmodes)]
    P, W = A.BATCH_PIPELINED, A.BATCH_WIDE
    for plan in A.plan_candidates(t.nmodes, t.dims):
        e = A.Engine(t, R, plan, precision=prec, tiles=(64, 0, 0))
        for m in range(t.nmodes):
            e.set_factor(m, U[m])
        first = None
        for batch in (0, 4, 8, 16, 4 | P, 8 | P, 4 | W, 8 | W):
            e.set_batch(batch)
            out = [e.mttkrp(m) for m in range(t.nmodes)]
            for m in range(t.nmodes):
                assert O.rel_fro(out[m], ref[m]) <= TOL[prec], (plan.groups(t.nmodes), batch, m)
            if first is None:
                first = out
            else:
                for m in range(t.nmodes):
                    assert np.array_equal(out[m], first[m]), (plan.groups(t.nmodes), batch, m)
        e.close()


@pytest.mark.parametrize("prec", [A.FP32, A.FP64, A.BF16ROWS])
def test_node_rows_buffer_and_global_paths_agree(prec):
    """The fused reductions read intermediate nodes' rows through buffer
    loads (a non-closing child loads nothing) when every factor spans less
    than 2^31 bytes, else through a stand-in row (ENGINE_GLOBAL_NODE_ROWS
    forces that path): every plan's MTTKRPs, leaf to root and staged, in the
    wide and narrow layouts, agree bit for bit, and with the oracle."""
    print(f"node rows prec={prec}", flush=True)
    t = _tensor([150, 100, 60, 40, 20], 40000, A.DIST_ZIPF, seed=43)
    R = 32
    U = _factors(t, R, seed=6)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(t.nmodes)]
    for plan in A.plan_candidates(t.nmodes, t.dims):
        for staged in (False, True):
            outs = []
            for fl in (0, A.ENGINE_GLOBAL_NODE_ROWS):
                e = A.Engine(t, R, plan, precision=prec, tiles=(32, 16, 0), staged=staged,
                             flags=fl)
                for m in range(t.nmodes):
                    e.set_factor(m, U[m])
                o = []
                for batch in (0, 8 | A.BATCH_WIDE, 4):
                    e.set_batch(batch)
                    o += [e.mttkrp(m) for m in range(t.nmodes)]
                outs.append(o)
                e.close()
            for m in range(t.nmodes):
                assert O.rel_fro(outs[0][m], ref[m]) <= TOL[prec], (plan.groups(t.nmodes), m)
            for a, b in zip(outs[0], outs[1]):
                assert np.array_equal(a, b), (plan.groups(t.nmodes), staged)


@pytest.mark.parametrize("precs", [(A.FP32,) * 4, (A.FP64,) * 4,
                                   (A.FP32, A.FP64, A.FP32, A.FP64)])
def test_batched_cpd_matches_each_instance(precs):
    """Uneven instances (orders 4-6, ranks 8/16/32) swept together in one
    hipGraph: every replay's fits equal each engine's own eager sweeps.  One
    precision runs the grouped dense updates (FP32: MFMA Grams; FP64: the
    FP64 Gram, product and fit), mixed precisions one branch per instance."""
    print(f"batched cpd precisions={precs}", flush=True)
    specs = [([60, 50, 40, 30], 8000, A.DIST_UNIFORM, 8),
             ([80, 70, 60, 50, 40], 12000, A.DIST_ZIPF, 16),
             ([40, 35, 30, 25, 20, 15], 6000, A.DIST_UNIFORM, 32),
             ([100, 90, 80, 70], 20000, A.DIST_ZIPF, 16)]
    eng, solo = [], []
    for k, (dims, nd, dist, R) in enumerate(specs):
        t = _tensor(dims, nd, dist, seed=60 + k)
        cands, ch = A.plan_analytic(t, R)
        for lst in (eng, solo):
            e = A.Engine(t, R, cands[ch], precision=precs[k])
            e.init_factors(7 + k)
            lst.append(e)
    b = A.Batch(eng)
    ms, fits = b.cpd(5)
    assert ms > 0
    for k, e in enumerate(solo):
        ref = np.asarray(e.cpd(5, tol=0.0, graph=False))
        assert np.array_equal(fits[k], ref), (k, fits[k], ref)
    # a second batched run from the updated factors replays the same graph
    ms2, fits2 = b.cpd(3)
    for k, e in enumerate(solo):
        ref = np.asarray(e.cpd(3, tol=0.0, graph=False))
        assert np.array_equal(fits2[k], ref), (k, fits2[k], ref)
    assert b.sequential(2) > 0
    b.close()


@pytest.mark.parametrize("form", [A.FORM_PERSISTENT, A.FORM_GROUPED])
@pytest.mark.parametrize("prec", [A.FP32, A.FP64])
def test_persistent_batch_equals_forked_branches(prec, form):
    """Each round's MTTKRPs of every instance recorded as stages and run in
    one persistent launch (blocks taken from per-XCD queues in dependency
    order) or as grouped launches (one per depth and kind of stage): fits,
    factors and lambdas bit-identical to the forked-branch batch, over two
    runs of the same captured graph; Zipf instances bring split parents
    (merge stages) and empty root slices (zero-fill stages)."""
    print(f"batch form={form} precision={prec}", flush=True)
    specs = [([60, 50, 40, 30], 8000, A.DIST_UNIFORM, 8),
             ([300, 70, 60, 50, 40], 30000, A.DIST_ZIPF, 16),
             ([40, 35, 30, 25, 20, 15], 6000, A.DIST_UNIFORM, 32),
             ([900, 90, 80, 70], 40000, A.DIST_ZIPF, 16),
             ([500, 400, 300, 200, 100, 50], 60000, A.DIST_ZIPF, 32),
             ([2000, 30, 20, 10], 5000, A.DIST_ZIPF, 24)]
    sets = []
    for _ in range(2):
        eng = []
        for k, (dims, nd, dist, R) in enumerate(specs):
            t = _tensor(dims, nd, dist, seed=90 + k)
            cands, ch = A.plan_analytic(t, R)
            e = A.Engine(t, R, cands[ch], precision=prec)
            e.init_factors(17 + k)
            eng.append(e)
        sets.append(eng)
    bp, bf = A.Batch(sets[0]), A.Batch(sets[1])
    bp.set_form(form)
    bf.set_form(A.FORM_FORKED)
    for n in (4, 3):
        _, fp = bp.cpd(n)
        _, ff = bf.cpd(n)
        assert bp.form == form and bf.form == A.FORM_FORKED
        assert bp.persistent == (form == A.FORM_PERSISTENT) and not bf.persistent
        assert np.array_equal(fp, ff), (fp, ff)
        for k, (ep, ef) in enumerate(zip(*sets)):
            assert np.array_equal(ep.get_lambda(), ef.get_lambda()), k
            for m in range(len(specs[k][0])):
                assert np.array_equal(ep.get_factor(m), ef.get_factor(m)), (k, m)
    bp.close()
    bf.close()
    # outside its range (R > 32) the batch keeps the forked branches
    t = _tensor([40, 30, 20, 10], 4000, A.DIST_UNIFORM, seed=99)
    cands, ch = A.plan_analytic(t, 40)
    e = A.Engine(t, 40, cands[ch], precision=prec)
    e.init_factors(5)
    b = A.Batch([e])
    with pytest.raises(A.AdatmError):
        b.set_form(form)
    b.cpd(1)
    assert b.form == A.FORM_FORKED
    b.close()


def test_device_fiber_counts_match_host():
    """Radix sort + boundary scan on the device give the host's (and the
    oracle's) fiber counts, including keys that need three 8-bit digits."""
    spec = A.make_spec([70000, 300, 5, 40], 300000, dist=A.DIST_ZIPF, zipf_s=0.7, seed=31)
    t = A.Tensor.synth(spec)
    t.canonicalize()
    inds = t.inds().astype(np.int64)
    for order in ([0, 1, 2, 3], [3, 2, 1, 0], [2, 0, 3, 1], [1, 3, 0, 2]):
        dev = A.count_fibers(t, order, device=0)
        assert dev == A.count_fibers(t, order, device=-1) == O.count_fibers(inds, order)


def test_device_csf_large_bit_exact():
    """Device-built CSFs of a 2M-draw tensor equal the oracle's structure."""
    t = A.synth_tensor(1, 2_000_000)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    cands, ch = A.plan_analytic(t, 16, device=0)
    e = A.Engine(t, 16, cands[ch], precision=A.FP64)
    for g, grp in enumerate(cands[ch].groups(4)):
        F, fptr, fids, gv = e.get_csf(g, t.nnz)
        Fo, fpo, fio, vo, _ = O.csf(inds, vals, grp[3])
        assert F == Fo
        for l in range(3):
            assert np.array_equal(fptr[l], fpo[l])
        for l in range(1, 4):
            assert np.array_equal(fids[l], fio[l])
        assert np.array_equal(gv, vo)
    U = [A.rand_matrix(4, m, d, 16) for m, d in enumerate(t.dims)]
    for m in range(4):
        e.set_factor(m, U[m])
    for m in range(4):
        assert O.rel_fro(e.mttkrp(m), O.coo_mttkrp(inds, vals, U, m)) <= 1e-12


def test_planner_device_equals_host():
    t = A.synth_tensor(2, 300_000)
    a, ca = A.plan_analytic(t, 32, device=0)
    b, cb = A.plan_analytic(t, 32, device=-1)
    assert ca == cb and [(p.pspace, p.pops) for p in a] == [(p.pspace, p.pops) for p in b]


def _degenerate_cases():
    rng = np.random.default_rng(77)
    out = []
    # one nonzero (every mode compacts to one index)
    out.append(("one nonzero", np.array([[3], [2], [1], [0]]), np.array([1.5]), [5, 4, 3, 2], 8))
    # one fiber: a single (i0, i1, i2) path with 3000 leaves (pieces split it)
    n = 3000
    out.append(("one fiber", np.stack([np.full(n, 1), np.full(n, 2), np.full(n, 0), np.arange(n)]),
                rng.standard_normal(n), [3, 4, 2, n], 16))
    # one root slice holding every nonzero
    n = 6000
    out.append(("one root slice", np.stack([np.zeros(n, dtype=np.int64), rng.integers(0, 40, n),
                                            rng.integers(0, 30, n), rng.integers(0, 20, n)]),
                rng.standard_normal(n), [1, 40, 30, 20], 32))
    # rank one, duplicates merged by canonicalisation
    n = 4000
    idx = np.stack([rng.integers(0, 15, n), rng.integers(0, 12, n), rng.integers(0, 10, n)])
    out.append(("rank one, duplicates", idx, rng.standard_normal(n), [15, 12, 10], 1))
    return out


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("prec", [A.FP64, A.FP32])
def test_degenerate_tensors(case, prec):
    """Edge shapes: a single nonzero, a single fiber split across pieces, a
    single root slice, rank one with duplicate coordinates: every plan and
    mode against the oracle's COO MTTKRP of the canonical tensor."""
    name, inds, vals, dims, R = _degenerate_cases()[case]
    print(f"degenerate {name} prec={prec}", flush=True)
    t = A.Tensor.from_arrays(inds.astype(np.uint32), vals.astype(np.float64), dims)
    t.canonicalize()
    U = _factors(t, R, seed=3)
    ci, cv = t.inds().astype(np.int64), t.vals.copy()
    ref = [O.coo_mttkrp(ci, cv, U, m) for m in range(t.nmodes)]
    for plan in A.plan_candidates(t.nmodes, t.dims):
        e = A.Engine(t, R, plan, precision=prec, tiles=(32, 16, 0))
        for m in range(t.nmodes):
            e.set_factor(m, U[m])
        for m in range(t.nmodes):
            err = O.rel_fro(e.mttkrp(m), ref[m])
            assert err <= TOL[prec], (name, plan.groups(t.nmodes), m, err)
        e.close()


@pytest.mark.parametrize("R", [4, 40])
def test_solve_status_reports_a_ridge(R):
    """Factors whose columns are all equal make every hadamard of Grams
    exactly singular (powers of two keep the elimination exact), so the first
    solve finds a zero pivot and solves with a ridge: the host reads status 1
    and the fits stay finite (R = 4: the one-wave elimination; R = 40: the
    four-wave one).  A well-posed run reads status 0."""
    print(f"solve status R={R}", flush=True)
    I, J, K = 8, 4, 16
    g = np.stack(np.meshgrid(np.arange(I), np.arange(J), np.arange(K), indexing="ij"))
    inds = g.reshape(3, -1)
    vals = np.random.default_rng(1).random(inds.shape[1]) + 0.5
    t = A.Tensor.from_arrays(inds, vals, [I, J, K])
    t.canonicalize()
    assert t.dims == [I, J, K]
    plan = A.plan_candidates(3, t.dims)[0]
    e = A.Engine(t, R, plan, precision=A.FP64)
    for m, d in enumerate(t.dims):
        e.set_factor(m, np.ones((d, R)))
    e.cpd_begin()
    fits = [e.cpd_sweep(it) for it in range(3)]
    assert e.solve_status() == 1
    assert np.all(np.isfinite(fits)), fits
    e.close()
    t2 = _tensor([50, 40, 30, 20], 20000, A.DIST_UNIFORM, seed=21)
    cands, ch = A.plan_analytic(t2, 8)
    e2 = A.Engine(t2, 8, cands[ch], precision=A.FP32)
    e2.init_factors(3)
    e2.cpd(4, graph=True)
    assert e2.solve_status() == 0


def test_stepwise_sweeps_follow_the_device_counter():
    """cpd_sweep returns the fit its own sweep wrote: the sweep index must be
    the number of sweeps since the start, CP-ALS runs continue the count, and
    no sweep runs once the fits buffer (4096) is full."""
    print("sweep counter", flush=True)
    t = _tensor([40, 30, 20, 10], 5000, A.DIST_UNIFORM, seed=33)
    cands, ch = A.plan_analytic(t, 8)
    e = A.Engine(t, 8, cands[ch], precision=A.FP64)
    e.init_factors(5)
    e.cpd_begin()
    with pytest.raises(A.AdatmError):
        e.cpd_sweep(1)
    f0 = e.cpd_sweep(0)
    f1 = e.cpd_sweep(1, graph=True)
    assert np.allclose(e.fits(2), [f0, f1], rtol=0, atol=0)
    e.init_factors(5)
    fits = e.cpd(3, graph=True)  # starts over: sweeps 0..2
    assert fits[0] == f0 and fits[1] == f1
    f3 = e.cpd_sweep(3)  # continues the run
    assert f3 == e.fits(4)[3]
    with pytest.raises(A.AdatmError):
        e.cpd_sweep(0)
    with pytest.raises(A.AdatmError):
        e.run_timed(1, 4096, graph=True)
    e.cpd_begin()
    for it in range(4096):
        e.cpd_sweep(it, graph=True)
    with pytest.raises(A.AdatmError):
        e.cpd_sweep(4096, graph=True)
    e.close()


@pytest.mark.parametrize("prec", [A.FP64, A.FP32])
def test_memo_levels_written_beside_the_root_pass(prec):
    """The memo levels the one-pass root MTTKRP writes (reuse modes read
    them) equal those of the staged chain: every mode of every plan agrees
    between the two engines (FP64 to rounding, FP32 within 1e-6), over
    tiles that cut long fibers into many pieces, and CP-ALS fits agree."""
    print(f"fused vs staged prec={prec}", flush=True)
    t = _tensor([300, 200, 100, 50, 20], 120000, A.DIST_ZIPF, seed=41)
    R = 16
    U = _factors(t, R, seed=8)
    for plan in A.plan_candidates(t.nmodes, t.dims):
        out = {}
        for staged, tiles in ((True, (0, 0)), (False, (8, 8)), (False, (64, 0)),
                              (False, (512, 0))):
            e = A.Engine(t, R, plan, precision=prec, staged=staged, tiles=tiles)
            for m in range(t.nmodes):
                e.set_factor(m, U[m])
            out[(staged, tiles)] = [e.mttkrp(m) for m in range(t.nmodes)]
            e.close()
        ref = out[(True, (0, 0))]
        for key, M in out.items():
            for m in range(t.nmodes):
                err = O.rel_fro(M[m], ref[m])
                assert err <= (1e-13 if prec == A.FP64 else 1e-6), (plan.groups(5), key, m, err)
