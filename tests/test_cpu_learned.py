"""Learned planner on the CPU: the oracle network pinned by mathematics,
and the plan features computed with host fiber counts."""
import math

import numpy as np
import pytest

import adatm_amd as A
from oracle import mlpref as M


def test_param_count_matches_library(lib):
    assert M.nparams() == 9985 == A.Mlp.nparams()


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.he_init(1)
    X = rng.normal(size=(13, 24))
    t = rng.normal(size=13)
    _, g = M.loss_grad(p, X, t)
    idx = rng.choice(p.size, 60, replace=False)
    # every bias of the last layer and a few weights of the first too
    idx = np.concatenate([idx, [p.size - 1, 0, 1, 24 * 64 + 3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (M.loss_grad(p + e, X, t)[0] - M.loss_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_oracle_forward_is_relu_mlp():
    """A network whose layers are identities passes ReLU(x) of the first
    input through; a dropped ReLU or transposed weight breaks this."""
    p = np.zeros(M.nparams())
    L = M.unpack(p)
    L[0][0][0, 0] = 1.0            # W1: input 0 -> hidden 0
    L[1][0][0, 0] = 2.0            # W2
    L[2][0][0, 0] = 1.0            # W3
    L[3][0][0, 0] = 3.0            # W4
    L[3][1][0] = 0.5               # b4
    X = np.zeros((3, 24))
    X[:, 0] = [-1.0, 0.0, 2.0]
    X[:, 1] = [5.0, 5.0, 5.0]      # unused input
    y, _ = M.forward(p, X)
    assert np.allclose(y, [0.5, 0.5, 0.5 + 6.0 * 2.0])


def test_oracle_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, _, _ = M.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-2)
    assert np.allclose(p1, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)


def _tensor(N=5, ndraws=20000, seed=11):
    dims = [60, 50, 40, 30, 20, 15, 12, 10][:N]
    spec = A.make_spec(dims, ndraws, dist=A.DIST_ZIPF, zipf_s=0.8, seed=seed)
    t = A.Tensor.synth(spec)
    t.canonicalize()
    return t


@pytest.mark.parametrize("N", [3, 4, 5, 6, 8])
def test_features_host(lib, N):
    t = _tensor(N)
    R = 16
    cands = A.plan_candidates(N, t.dims)
    f = A.plan_features(t, R, cands, device=-1)
    assert f.shape == (len(cands), 24) and np.isfinite(f).all()
    assert (f[:, 0] == N).all()
    assert np.allclose(f[:, 1], math.log10(t.nnz), atol=1e-6)
    assert np.allclose(f[:, 2], math.log2(R))
    assert (f[:, 7] == [c.n_grp for c in cands]).all()
    # feature 12 is the reference's op model on fresh statistics
    ref, _ = A.plan_analytic(t, R, ref_quirks=False, device=-1)
    pops = np.array([c.pops for c in ref], dtype=np.float64)
    assert np.allclose(f[:, 12], np.log10(pops), rtol=0, atol=1e-5)
    # deterministic
    assert np.array_equal(f, A.plan_features(t, R, cands, device=-1))
    # the same features from counts given explicitly
    for c, plan in enumerate(cands):
        Fg = [A.count_fibers(t, plan.grp[g].o_opt[:N]) for g in range(plan.n_grp)]
        fc = A.plan_features_from_counts(N, t.nnz, t.dims, R, plan, Fg)
        assert np.array_equal(fc, f[c])


def test_committed_corpus_is_consistent():
    """The committed planner corpus: every tensor's candidates are complete
    (Appendix A counts for its order), timings positive, the analytic choice
    is a valid candidate, and the held-out split has 80 tensors."""
    import glob
    import os
    from adatm_amd import corpus as K
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    paths = sorted(glob.glob(os.path.join(root, "adatm_amd", "planner", "corpus_c4_*.npz")))
    if not paths:
        pytest.skip("no committed corpus")
    d = K.load(paths)
    counts = {3: 1, 4: 3, 5: 4, 6: 7, 7: 8, 8: 11}
    assert (d["ms"] > 0).all() and np.isfinite(d["X"]).all()
    for k, N, nnz, R, nc, an in d["meta"]:
        assert nc == counts[int(N)] == (d["tensor"] == k).sum()
        assert 0 <= an < nc and R in (8, 16, 32, 64)
        spec, Rs = K.corpus_spec(int(k))
        assert spec.nmodes == N and Rs == R
    assert len(np.unique(d["tensor"])) == K.NTENSORS
    assert len(np.unique(K.split(d, False)["tensor"])) == K.NTENSORS - K.NTRAIN
