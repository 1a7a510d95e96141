"""The reference drivers' call sequences, compiled as C against
include/splatt.h (tests/boundary_replay.c, built by build()) and run on the
GPU: bench_adaptm (src/adatm_bench.c:531-650), splatt_cpd_cmd_adaptive
(src/cmds/cmd_cpd.c:132-259), cpd_als_iterate_adaptive on caller matrices,
splatt_cpd_cmd (splatt_csf_alloc + splatt_cpd_als) and splatt_csf_load +
splatt_mttkrp.  Every output is checked against the oracle: the plan and use
map against the Python restatement of the planner, fiber counts against
brute force, CSF/vCSF storage against its closed form, MTTKRPs against the
COO brute force (<= 1e-12, FP64), CP-ALS fits (<= 1e-9) and the
post-processed factors and lambda (<= 1e-6) against the oracle's ALS."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "adatm_amd" / "bin" / "boundary_replay"
TNS = ROOT / "tests" / "golden" / "small4.tns"
FAC_SEED, ITER_SEED, CPD_SEED = 7, 9, 5
CSF_SPARSITY_BYTES = 88 * 8 * 3 + 8  # nfibs, fptr, fids arrays of 88 + vals pointer


def _read_tns(path):
    a = np.loadtxt(path, comments="#", ndmin=2)
    N = a.shape[1] - 1
    inds = a[:, :N].T.astype(np.int64) - 1
    return inds, a[:, N].copy(), [int(x) + 1 for x in inds.max(axis=1)]


def _kruskal(path, dims, R):
    v = np.fromfile(path, dtype=np.float64)
    fit, lam, off, fac = v[0], v[1:1 + R], 1 + R, []
    for d in dims:
        fac.append(v[off:off + d * R].reshape(d, R))
        off += d * R
    assert off == v.size
    return fit, lam, fac


def _post_process(A_, lam):
    A_ = [a.copy() for a in A_]
    lam = lam.copy()
    for a in A_:
        s = np.sqrt((a * a).sum(axis=0))
        a /= np.where(s > 0, s, 1.0)
        lam *= s
    return A_, lam


def _check_cpd(path, inds, vals, dims, R, niters, seed):
    fit, lam, fac = _kruskal(path, dims, R)
    U = [A.rand_matrix(seed, m, d, R) for m, d in enumerate(dims)]
    ofits, oA, olam = O.cpd_als(lambda m, F: O.coo_mttkrp(inds, vals, F, m), U, niters,
                                float((vals ** 2).sum()))
    assert abs(fit - ofits[-1]) <= 1e-9, (path.name, fit, ofits[-1])
    pA, plam = _post_process(oA, olam)
    assert O.rel_fro(lam, plam) <= 1e-6, path.name
    for m in range(len(dims)):
        assert O.rel_fro(fac[m], pA[m]) <= 1e-6, (path.name, m)


CASES = [("fixture", str(TNS), 8, 10), ("C1 prefix", "synth:1:200000", 16, 5)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_reference_call_sequences(case, tmp_path):
    name, src, R, niters = CASES[case]
    print(f"boundary replay: {name} R={R} iters={niters}", flush=True)
    assert BIN.exists(), "build() compiles tests/boundary_replay.c"
    r = subprocess.run([str(BIN), src, str(R), str(niters), str(tmp_path), str(TNS)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    if src.startswith("synth:"):
        t = A.synth_tensor(1, 200_000)
        inds, vals, dims = t.inds().astype(np.int64), t.vals.copy(), t.dims
    else:
        inds, vals, dims = _read_tns(src)
    N, nnz = len(dims), vals.size

    # plan, use map, fiber counts, storage
    info = {"group": [], "nfibs": []}
    for line in (tmp_path / "plan.txt").read_text().splitlines():
        k, *rest = line.split()
        if k in ("group", "nfibs"):
            info[k].append([int(x) for x in rest])
        elif k == "use":
            info[k] = [tuple(int(y) for y in x.split(":")) for x in rest]
        else:
            info[k] = int(rest[0])
    configs = O.enumerate_plans(N, dims)
    F_of = {}

    def fibers(o):
        if o not in F_of:
            F_of[o] = O.count_fibers(inds, list(o))
        return F_of[o]
    ch = O.choose(O.predict(configs, fibers, nnz, N, R))
    groups = configs[ch]
    assert [(g[0], g[1], g[2], *g[3]) for g in groups] == [tuple(x) for x in info["group"]]
    grp, tag = O.plan_use(groups, N)
    assert info["use"] == list(zip(grp, tag))
    assert info["n_csf"] == len(groups)
    assert info["n_rcsf"] == sum(1 for g in groups if g[2] > 0)
    cs_bytes = rcsf_bytes = 0
    for g, (mm, b, nim, o) in enumerate(groups):
        F = fibers(tuple(o))
        assert info["nfibs"][g] == [g] + list(F)
        # vals + leaf ids + the sparsity struct + fptr and ids of levels < N-1
        # (root ids kept: fids[0] is not NULL here)
        cs_bytes += nnz * 8 + nnz * 8 + CSF_SPARSITY_BYTES
        cs_bytes += sum((F[l] + 1) * 8 + F[l] * 8 for l in range(N - 1))
        for k in range(nim):  # vCSF k: F[b-1-k] rows of R values
            rcsf_bytes += F[b - 1 - k] * R * 8 + CSF_SPARSITY_BYTES
    assert info["cs_bytes"] == cs_bytes and info["rcsf_bytes"] == rcsf_bytes

    # MTTKRP sequence (mttkrp_csf_adaptive per mode)
    U = [A.rand_matrix(FAC_SEED, m, d, R) for m, d in enumerate(dims)]
    for m in range(N):
        M = np.fromfile(tmp_path / f"mttkrp_{m}.bin").reshape(dims[m], R)
        assert O.rel_fro(M, O.coo_mttkrp(inds, vals, U, m)) <= 1e-12, m

    # the three CP-ALS entry points
    _check_cpd(tmp_path / "cpd_adaptive.bin", inds, vals, dims, R, niters, CPD_SEED)
    _check_cpd(tmp_path / "iterate.bin", inds, vals, dims, R, niters, ITER_SEED)
    _check_cpd(tmp_path / "cpd_plain.bin", inds, vals, dims, R, niters, CPD_SEED)

    # splatt_csf_load + splatt_mttkrp on the fixture file
    li, lv, ld = _read_tns(TNS)
    LU = [A.rand_matrix(FAC_SEED, m, d, R) for m, d in enumerate(ld)]
    for m in range(len(ld)):
        M = np.fromfile(tmp_path / f"load_mttkrp_{m}.bin").reshape(ld[m], R)
        assert O.rel_fro(M, O.coo_mttkrp(li, lv, LU, m)) <= 1e-12, m
