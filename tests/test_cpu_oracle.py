"""Pins of the oracle itself (CPU only).

The oracle is checked against what the mathematics fixes, chosen so that a
dropped term, a wrong index or a transposed operand fails:
  - brute-force COO MTTKRP == dense unfolding times the Khatri-Rao product
    built independently with numpy (einsum on a densified tiny tensor);
  - the memoized sequence by definition (npref.MemoSeq) and the C/OpenMP
    restatement (corc) == brute force on every mode of every candidate plan;
  - one-fiber and one-slice tensors reduce to an R-wide axpy / dot;
  - CP-ALS: the fit of the oracle equals 1 - ||X - Z||/||X|| computed on the
    densified tensor and the reconstructed Kruskal tensor;
  - Philox4x32-10 against the Random123 known-answer vectors.
"""
import numpy as np
import pytest

import adatm_amd as A
from oracle import corc
from oracle import npref as O


def _rand_tensor(rng, dims, nnz):
    inds = np.stack([rng.integers(0, d, nnz) for d in dims]).astype(np.int64)
    vals = rng.random(nnz)
    ui, uv, nd, _ = O.canonicalize(inds, vals, dims)
    return ui, uv, nd


def _dense_mttkrp(X, U, mode):
    """X_(mode) (KR of the other factors), via an einsum over the dense tensor."""
    N = X.ndim
    letters = "abcdefgh"[:N]
    spec = letters + "," + ",".join(f"{letters[k]}r" for k in range(N) if k != mode)
    spec += f"->{letters[mode]}r"
    return np.einsum(spec, X, *[U[k] for k in range(N) if k != mode])


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert A.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert A.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert A.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                    [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("N", [3, 4, 5])
def test_coo_mttkrp_equals_dense_unfolding(N):
    rng = np.random.default_rng(N)
    dims = [5, 4, 6, 3, 4][:N]
    ui, uv, nd = _rand_tensor(rng, dims, 60)
    X = O.dense_tensor(ui, uv, nd)
    U = [rng.random((d, 3)) for d in nd]
    for m in range(N):
        np.testing.assert_allclose(O.coo_mttkrp(ui, uv, U, m), _dense_mttkrp(X, U, m),
                                   rtol=1e-13, atol=1e-13)


def test_coo_mttkrp_catches_planted_mistake():
    """A transposed operand (wrong factor for a mode) is detected."""
    rng = np.random.default_rng(7)
    ui, uv, nd = _rand_tensor(rng, [5, 5, 5], 50)
    X = O.dense_tensor(ui, uv, nd)
    U = [rng.random((d, 2)) for d in nd]
    wrong = [U[0], U[2], U[1]]
    assert not np.allclose(O.coo_mttkrp(ui, uv, wrong, 0), _dense_mttkrp(X, U, 0))


def test_single_fiber_is_axpy():
    rng = np.random.default_rng(1)
    nnz = 7
    inds = np.stack([np.zeros(nnz, np.int64), np.zeros(nnz, np.int64), np.arange(nnz)])
    vals = rng.random(nnz)
    U = [rng.random((1, 4)), rng.random((1, 4)), rng.random((nnz, 4))]
    M0 = O.coo_mttkrp(inds, vals, U, 0)
    np.testing.assert_allclose(M0[0], U[1][0] * (vals @ U[2]), rtol=1e-14)
    M2 = O.coo_mttkrp(inds, vals, U, 2)
    np.testing.assert_allclose(M2, vals[:, None] * (U[0][0] * U[1][0])[None, :], rtol=1e-14)


CASES = [([12, 10, 8], 300), ([9, 8, 7, 6], 800), ([7, 6, 6, 5, 4], 1500),
         ([5, 5, 4, 4, 4, 3], 1500), ([4] * 7, 2500), ([3, 4, 3, 4, 3, 4, 3, 4], 3000)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_memo_sequences_match_brute_force(case):
    dims, nnz = CASES[case]
    rng = np.random.default_rng(100 + case)
    ui, uv, nd = _rand_tensor(rng, dims, nnz)
    R = 5
    U = [rng.random((d, R)) for d in nd]
    ref = [O.coo_mttkrp(ui, uv, U, m) for m in range(len(nd))]
    for plan in O.enumerate_plans(len(nd), nd):
        ms = O.MemoSeq(ui, uv, nd, plan)
        cs = corc.OracleSeq(ui, uv, nd, plan, R)
        for m in range(len(nd)):
            assert O.rel_fro(ms.mttkrp(m, U), ref[m]) < 1e-13, (plan, m)
            assert O.rel_fro(cs.mttkrp(m, U), ref[m]) < 1e-13, (plan, m)


def test_oracle_cpd_fit_equals_dense_residual():
    rng = np.random.default_rng(11)
    ui, uv, nd = _rand_tensor(rng, [6, 5, 4, 5], 400)
    X = O.dense_tensor(ui, uv, nd)
    R = 3
    U = [rng.random((d, R)) for d in nd]
    fits, Af, lam = O.cpd_als(lambda m, F: O.coo_mttkrp(ui, uv, F, m), U, 4,
                              float((uv ** 2).sum()))
    Z = np.einsum("r,ar,br,cr,dr->abcd", lam, *Af)
    direct = 1.0 - np.linalg.norm(X - Z) / np.linalg.norm(X)
    assert abs(fits[-1] - direct) < 1e-10


def test_oracle_cpd_c_equals_numpy():
    rng = np.random.default_rng(12)
    ui, uv, nd = _rand_tensor(rng, [20, 15, 10, 8], 1500)
    R = 6
    U = [rng.random((d, R)) for d in nd]
    plans = O.enumerate_plans(4, nd)
    ref, _, lam_ref = O.cpd_als(lambda m, F: O.coo_mttkrp(ui, uv, F, m), U, 6,
                                float((uv ** 2).sum()))
    for plan in plans:
        f, _, lam = corc.OracleSeq(ui, uv, nd, plan, R).cpd(U, 6)
        assert np.max(np.abs(f - np.asarray(ref))) < 1e-12
        np.testing.assert_allclose(lam, lam_ref, rtol=1e-10)


def test_als_update_solves_normal_equations():
    """One ALS update A = M V^-1 makes the gradient of ||X - Z||^2 in A vanish."""
    rng = np.random.default_rng(13)
    ui, uv, nd = _rand_tensor(rng, [6, 5, 4], 100)
    X = O.dense_tensor(ui, uv, nd)
    U = [rng.random((d, 3)) for d in nd]
    M = O.coo_mttkrp(ui, uv, U, 0)
    V = (U[1].T @ U[1]) * (U[2].T @ U[2])
    A0 = M @ np.linalg.inv(V)
    Z = np.einsum("ar,br,cr->abc", A0, U[1], U[2])
    grad = _dense_mttkrp(Z - X, [A0, U[1], U[2]], 0)
    assert np.max(np.abs(grad)) < 1e-10
