"""GPU parity on the named workloads' own generators (the benchmark
configurations), against the FP64 oracle.

* 4-mode 2000x1500x1000x500, 2M uniform draws, R=16: every mode's MTTKRP
  and 20 CP-ALS sweeps at full size, FP64 and FP32.
* 5-mode 1e5x5e4x2e4x5e3x1e3, Zipf(1.1), R=32: the first 2M draws against
  the brute-force oracle (every mode, FP64 and FP32) and the oracle's CP-ALS;
  the full 100M draws through size-independent properties (FP32 against FP64
  on the device: MTTKRP and fits).
* 8-mode 1000^8, Gaussian clusters, R=64: the first 1M draws, the analytic
  plan (8) and the learned one (10), FP32 and BF16 factor rows.
* 64 uneven batched instances: the first 8, batched in one graph, 50 sweeps,
  fits against the oracle's.
Tolerances: MTTKRP relative Frobenius <= 1e-12 (FP64), 1e-5 (FP32), 2e-2
(BF16 rows); fits within 1e-9 (FP64), 1e-4 (FP32) after the same sweeps.
"""
import numpy as np
import pytest

import adatm_amd as A
from oracle import corc
from oracle import npref as O

pytestmark = pytest.mark.gpu

TOL = {A.FP64: 1e-12, A.FP32: 1e-5, A.BF16ROWS: 2e-2}
FIT_TOL = {A.FP64: 1e-9, A.FP32: 1e-4}


def _factors(t, R, seed):
    return [A.rand_matrix(seed, m, d, R) for m, d in enumerate(t.dims)]


def _engine(t, R, plan, prec, U):
    e = A.Engine(t, R, plan, precision=prec)
    for m in range(t.nmodes):
        e.set_factor(m, U[m])
    return e


def test_c1_full_mttkrp_and_20_sweeps():
    t = A.synth_tensor(1)
    R = 16
    print(f"C1 full: {t.nnz} nnz dims {t.dims} R={R}", flush=True)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    cands, ch = A.plan_analytic(t, R, device=0)
    assert ch == 2  # the reference model's choice at this workload
    U = _factors(t, R, seed=A.named_spec(1).seed + 1)
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(4)]
    seq = corc.OracleSeq(inds, vals, t.dims, cands[ch].groups(4), R)
    ofits, _, _ = seq.cpd(U, 20)
    seq.close()
    for prec in (A.FP64, A.FP32):
        e = _engine(t, R, cands[ch], prec, U)
        for m in range(4):
            assert O.rel_fro(e.mttkrp(m), ref[m]) <= TOL[prec], (prec, m)
        fits = e.cpd(20, tol=0.0, graph=True)
        assert np.max(np.abs(fits - ofits)) <= FIT_TOL[prec], (prec, fits, ofits)
        assert e.solve_status() == 0
        e.close()


def test_c2_prefix_against_the_oracle():
    t = A.synth_tensor(2, 2_000_000)
    R = 32
    print(f"C2 prefix: {t.nnz} nnz dims {t.dims} R={R}", flush=True)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    cands, ch = A.plan_analytic(t, R, device=0)
    assert ch == 3
    plan = cands[ch]
    U = _factors(t, R, seed=A.named_spec(2).seed + 1)
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(5)]
    seq = corc.OracleSeq(inds, vals, t.dims, plan.groups(5), R)
    ofits, _, _ = seq.cpd(U, 5)
    seq.close()
    for prec in (A.FP64, A.FP32):
        e = _engine(t, R, plan, prec, U)
        for m in range(5):
            err = O.rel_fro(e.mttkrp(m), ref[m])
            assert err <= TOL[prec], (prec, m, err)
        fits = e.cpd(5, tol=0.0, graph=True)
        assert np.max(np.abs(fits - ofits)) <= FIT_TOL[prec], (prec, fits, ofits)
        e.close()


def test_c2_full_size_fp32_against_fp64():
    t = A.synth_tensor(2)
    R = 32
    print(f"C2 full: {t.nnz} nnz dims {t.dims} R={R}", flush=True)
    cands, ch = A.plan_analytic(t, R, device=0)
    assert ch == 3
    U = _factors(t, R, seed=A.named_spec(2).seed + 1)
    out = {}
    fits = {}
    for prec in (A.FP64, A.FP32):
        e = _engine(t, R, cands[ch], prec, U)
        out[prec] = [e.mttkrp(m) for m in range(5)]
        fits[prec] = e.cpd(5, tol=0.0, graph=True)
        assert e.solve_status() == 0
        e.close()
    for m in range(5):
        err = O.rel_fro(out[A.FP32][m], out[A.FP64][m])
        assert err <= TOL[A.FP32], (m, err)
    assert np.max(np.abs(fits[A.FP32] - fits[A.FP64])) <= FIT_TOL[A.FP32], fits


@pytest.mark.parametrize("plan_idx", [8, 10])
def test_c3_prefix_fp32_and_bf16_rows(plan_idx):
    t = A.synth_tensor(3, 1_000_000)
    R = 64
    print(f"C3 prefix: {t.nnz} nnz R={R} plan {plan_idx}", flush=True)
    inds, vals = t.inds().astype(np.int64), t.vals.copy()
    plan = A.plan_candidates(8, t.dims)[plan_idx]
    U = _factors(t, R, seed=A.named_spec(3).seed + 1)
    ref = [O.coo_mttkrp(inds, vals, U, m) for m in range(8)]
    for prec in (A.FP32, A.BF16ROWS):
        e = _engine(t, R, plan, prec, U)
        for m in range(8):
            err = O.rel_fro(e.mttkrp(m), ref[m])
            assert err <= TOL[prec], (prec, m, err)
        e.close()


def test_c5_first_instances_batched_against_the_oracle():
    import bench
    n, iters = 8, 50
    print(f"C5: first {n} instances, {iters} batched sweeps", flush=True)
    eng, ofits, Us = [], [], []
    for k in range(n):
        spec, R = bench.c5_spec(k)
        t = A.Tensor.synth(spec)
        t.canonicalize()
        cands, ch = A.plan_analytic(t, R, device=0)
        U = _factors(t, R, seed=spec.seed + 1)
        Us.append(U)
        eng.append(_engine(t, R, cands[ch], A.FP32, U))
        seq = corc.OracleSeq(t.inds().astype(np.int64), t.vals, t.dims,
                             cands[ch].groups(t.nmodes), R)
        ofits.append(seq.cpd(U, iters)[0])
        seq.close()
    b = A.Batch(eng)
    # every batched form: one persistent launch per round, grouped launches,
    # forked branches
    for form in (A.FORM_PERSISTENT, A.FORM_GROUPED, A.FORM_FORKED):
        print(f"C5 batch form={form}", flush=True)
        for e, U in zip(eng, Us):
            for m, Um in enumerate(U):
                e.set_factor(m, Um)
        b.set_form(form)
        ms, fits = b.cpd(iters)
        assert ms > 0 and b.form == form
        for k in range(n):
            d = np.max(np.abs(fits[k] - ofits[k]))
            assert d <= FIT_TOL[A.FP32], (form, k, d)
    b.close()
