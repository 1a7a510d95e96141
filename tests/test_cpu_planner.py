"""Planner known answers (CPU only).

The expected plans below were derived by hand from src/adatm_bench.c:276-458
and src/adatm_sort.c:10-67 (candidate groups per order N, level orders at the
three named workloads, quicksort tie permutations) -- independent of both the
library's C planner and the oracle's Python transliteration, which are
checked against them and against each other.
"""
import math

import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

# (memo_mode, begin_imten, n_imten) per group, candidates in order
APPENDIX_A = {
    3: [[(0, 2, 1), (2, 1, 0)]],
    4: [[(0, 3, 2), (3, 1, 0)], [(0, 3, 1), (3, 1, 0)], [(0, 2, 1), (2, 2, 1)]],
    5: [[(0, 4, 3), (4, 1, 0)], [(0, 4, 2), (4, 1, 0)], [(0, 4, 1), (4, 1, 0)],
        [(0, 3, 2), (3, 2, 1)]],
    6: [[(0, 5, k), (5, 1, 0)] for k in (4, 3, 2, 1)] +
       [[(0, 3, 2), (3, 3, 2)], [(0, 3, 1), (3, 3, 1)],
        [(0, 2, 1), (2, 2, 1), (4, 2, 1)]],
    7: [[(0, 6, k), (6, 1, 0)] for k in (5, 4, 3, 2, 1)] +
       [[(0, 4, 3), (4, 3, 2)], [(0, 4, 2), (4, 3, 1)],
        [(0, 3, 2), (3, 2, 1), (5, 2, 1)]],
    8: [[(0, 7, k), (7, 1, 0)] for k in (6, 5, 4, 3, 2, 1)] +
       [[(0, 4, k), (4, 4, k)] for k in (3, 2, 1)] +
       [[(0, 3, 2), (3, 3, 2), (6, 2, 1)],
        [(0, 2, 1), (2, 2, 1), (4, 2, 1), (6, 2, 1)]],
}

# level orders at the named workloads' dims
ORDERS_C1 = {0: [[0, 1, 2, 3], [3, 2, 1, 0]], 1: [[0, 1, 2, 3], [3, 2, 1, 0]],
             2: [[0, 1, 3, 2], [2, 3, 1, 0]]}
ORDERS_C2 = {c: [[0, 1, 2, 3, 4], [4, 3, 2, 1, 0]] for c in range(3)}
ORDERS_C2[3] = [[0, 1, 2, 4, 3], [3, 4, 2, 1, 0]]
ORDERS_C3 = {c: [[0, 1, 2, 3, 4, 5, 6, 7], [7, 4, 6, 1, 5, 2, 0, 3]] for c in range(6)}
ORDERS_C3.update({c: [[0, 1, 2, 3, 4, 7, 5, 6], [4, 5, 6, 7, 0, 3, 1, 2]] for c in (6, 7, 8)})
ORDERS_C3[9] = [[0, 1, 2, 7, 6, 4, 3, 5], [3, 4, 5, 2, 1, 7, 6, 0], [6, 7, 4, 0, 1, 5, 2, 3]]
ORDERS_C3[10] = [[0, 1, 6, 2, 3, 7, 4, 5], [2, 3, 0, 4, 5, 1, 6, 7],
                 [4, 5, 2, 6, 7, 3, 0, 1], [6, 7, 4, 0, 1, 5, 2, 3]]

# pair_sort output when every key ties: position permutation per length
APPENDIX_B = {1: [0], 2: [0, 1], 3: [2, 0, 1], 4: [0, 3, 1, 2], 5: [4, 3, 1, 0, 2],
              6: [4, 0, 1, 5, 2, 3], 7: [4, 6, 1, 5, 2, 0, 3]}


@pytest.mark.parametrize("N", sorted(APPENDIX_A))
def test_candidate_groups(N):
    dims = [1000 + 10 * m for m in range(N)]
    got = [[(g[0], g[1], g[2]) for g in p.groups(N)] for p in A.plan_candidates(N, dims)]
    assert got == APPENDIX_A[N]
    ora = [[(g[0], g[1], g[2]) for g in p] for p in O.enumerate_plans(N, dims)]
    assert ora == APPENDIX_A[N]


@pytest.mark.parametrize("dims,orders", [
    ([2000, 1500, 1000, 500], ORDERS_C1),
    ([100000, 50000, 20000, 5000, 1000], ORDERS_C2),
    ([1000] * 8, ORDERS_C3),
])
def test_level_orders_at_named_workloads(dims, orders):
    N = len(dims)
    cands = A.plan_candidates(N, dims)
    for c, exp in orders.items():
        assert [g[3] for g in cands[c].groups(N)] == exp, c
    ora = O.enumerate_plans(N, dims)
    for c, exp in orders.items():
        assert [list(g[3]) for g in ora[c]] == exp, c


@pytest.mark.parametrize("n", sorted(APPENDIX_B))
def test_pair_sort_all_ties(n):
    pairs = [[p, 7] for p in range(n)]
    assert [x for x, _ in O.pair_sort(pairs)] == APPENDIX_B[n]
    arr = (A.C.c_uint64 * (2 * n))()
    for p in range(n):
        arr[2 * p], arr[2 * p + 1] = p, 7
    A.lib().pair_sort(arr, n)
    assert [arr[2 * i] for i in range(n)] == APPENDIX_B[n]


def test_pair_sort_random_ties_library_vs_transliteration():
    rng = np.random.default_rng(0)
    L = A.lib()
    for _ in range(500):
        n = int(rng.integers(1, 12))
        ys = rng.integers(0, 4, n)
        pairs = [[i, int(y)] for i, y in enumerate(ys)]
        arr = (A.C.c_uint64 * (2 * n))()
        for i, (x, y) in enumerate(pairs):
            arr[2 * i], arr[2 * i + 1] = x, y
        L.pair_sort(arr, n)
        got = [[arr[2 * i], arr[2 * i + 1]] for i in range(n)]
        assert got == O.pair_sort(pairs)
        assert [y for _, y in got] == sorted(ys.tolist())


def test_candidate_counts_formula():
    """n_grp_cal = ceil(N / sqrt(2(N-1))) and the counts 1,3,4,7,8,11."""
    for N, cnt in zip(range(3, 9), [1, 3, 4, 7, 8, 11]):
        assert len(A.plan_candidates(N, [100] * N)) == cnt
        assert math.ceil(N / math.sqrt(2 * (N - 1))) == (2 if N <= 6 else 3)


def _small(seed, dims, nnz, dist):
    spec = A.make_spec(dims, nnz, dist=dist, zipf_s=1.1, nclusters=40, sigma=2.0, seed=seed)
    t = A.Tensor.synth(spec)
    t.canonicalize()
    return t


@pytest.mark.parametrize("quirks", [True, False])
@pytest.mark.parametrize("case", [([30, 25, 20, 15], 5000, 0), ([40, 30, 20, 10, 8], 8000, 1),
                                  ([12] * 6, 6000, 2), ([8] * 8, 9000, 2)])
def test_analytic_model_matches_oracle(case, quirks):
    dims, nnz, dist = case
    t = _small(3, dims, nnz, dist)
    inds = t.inds().astype(np.int64)
    R = 16
    cands, ch = A.plan_analytic(t, R, ref_quirks=quirks)
    ocfg = O.enumerate_plans(t.nmodes, t.dims)
    pred = O.predict(ocfg, lambda o: O.count_fibers(inds, o), t.nnz, t.nmodes, R,
                     quirks=quirks)
    assert [(p.pspace, p.pops) for p in cands] == pred
    assert ch == O.choose(pred)


def test_fiber_counts_brute_force():
    rng = np.random.default_rng(5)
    t = _small(9, [9, 7, 5, 4], 600, 0)
    inds = t.inds()
    for order in ([0, 1, 2, 3], [3, 1, 0, 2], [2, 3, 1, 0]):
        F = A.count_fibers(t, order)
        brute = [len({tuple(inds[list(order[:l + 1]), j]) for j in range(t.nnz)})
                 for l in range(4)]
        assert F == brute == O.count_fibers(inds, order)
    del rng


def test_c1_analytic_plan_is_candidate_2():
    """At C1's dims the model predicts pops in ratio 13.41 : 13.94 : 10.93
    (closed form for uniform coordinates) and picks candidate 2."""
    t = A.synth_tensor(1, 400_000)
    cands, ch = A.plan_analytic(t, 16)
    assert ch == 2
    assert [g[3] for g in cands[2].groups(4)] == ORDERS_C1[2]


def test_plan_use_map():
    p = A.plan_candidates(8, [1000] * 8)[9]  # 3 groups, memo modes 0, 3, 6
    grp, tag = A.plan_use(p, 8)
    assert grp == [0, 0, 0, 1, 1, 1, 2, 2]
    assert tag == [0, 1, 1, 0, 1, 1, 0, 1]


def _mkplan(groups, N):
    p = A.Plan()
    p.n_grp = len(groups)
    for g, (mm, b, nim) in enumerate(groups):
        p.grp[g].memo_mode, p.grp[g].begin_imten, p.grp[g].n_imten = mm, b, nim
        for l in range(N):
            p.grp[g].o_opt[l] = (mm + l) % N
    return p


@pytest.mark.parametrize("groups,N,kept,grp,tag", [
    # single-group branch (Appendix A N=5 c0): the trailing group stores
    # nothing, so mode 4 comes from the leaves of group 0's CSF
    ([(0, 4, 3), (4, 1, 0)], 5, [0], [0, 0, 0, 0, 0], [0, 1, 1, 1, 2]),
    # every group memoises (N=5 c3): the same maps as TIME_EFFICIENT
    ([(0, 3, 2), (3, 2, 1)], 5, [0, 3], [0, 0, 0, 1, 1], [0, 1, 1, 0, 1]),
    # a group in the middle that stores nothing uses the CSF before it
    ([(0, 3, 1), (2, 1, 0), (3, 2, 1)], 5, [0, 3], [0, 0, 0, 1, 1], [0, 1, 2, 0, 1]),
    # one group: one CSF whatever it stores (adatm_rcsf.c:513-516)
    ([(0, 1, 0)], 4, [0], [0, 0, 0, 0], [0, 1, 1, 1]),
])
def test_space_efficient_plans(lib, groups, N, kept, grp, tag):
    q, g, t = A.plan_space_efficient(_mkplan(groups, N), N)
    assert [q.grp[i].memo_mode for i in range(q.n_grp)] == kept
    assert g == grp and t == tag


# ---------------------------------------------------------------------------
# Hand-derived known answers of the analytic model (adatm_bench.c:145-272,
# adatm_rcsf.c:329-406) on two tiny tensors, rank 2 (8-byte values and ids).
# Per-level fiber counts F_l (distinct length-(l+1) prefixes; F_{N-1} = nnz)
# are counted by hand for every level order the candidates use.  With those:
#   csf_bytes(F) = 16 nnz + sum_{l<N-1} (16 F_l + 8)
#   csf_ops(F)   = 2R sum_l F_l
#   rcsf_bytes   = sum_{i<n_imten} F[b - i] R 8   (D6: the reference reads
#                  level b-i; the memo it allocates is level b-1-i)
#   rcsf_ops     = sum_{i=1..nre} 2R (F[end-1] if i < end else F[i]),
#                  end = b - n_imten + 1, nre = next memo mode - memo mode - 1
# D5: only the first candidate of a run of equal group counts recounts; the
# later ones keep its csf_bytes/csf_ops and read F of its LAST group.  Every
# candidate at N = 4 and 5 has two groups, so all but the first are stale.
KA_N4 = dict(
    dims=[5, 4, 3, 2],
    coords=[(0, 0, 0, 0), (0, 0, 1, 1), (0, 1, 1, 0), (1, 0, 0, 0), (1, 0, 0, 1),
            (2, 3, 2, 1), (4, 1, 0, 0)],
    # F by order: [0,1,2,3] (4,5,6,7); [3,2,1,0] (2,5,6,7);
    #             [0,1,3,2] (4,5,7,7); [2,3,1,0] (3,5,6,7)
    fibers={(0, 1, 2, 3): [4, 5, 6, 7], (3, 2, 1, 0): [2, 5, 6, 7],
            (0, 1, 3, 2): [4, 5, 7, 7], (2, 3, 1, 0): [3, 5, 6, 7]},
    # c0 [(0,3,2) o=[0,1,2,3], (3,1,0) o=[3,2,1,0]]: csf 376+344, ops 88+80;
    #    rcsf bytes (F[3]+F[2]) 16 = 208; ops end=2: F[1]4 + F[2]4 = 20+24
    # c1 [(0,3,1),(3,1,0)] stale, F of [3,2,1,0]: bytes F[3]16 = 112;
    #    ops end=3: F[2]4 twice = 48
    # c2 [(0,2,1) o=[0,1,3,2], (2,2,1) o=[2,3,1,0]] stale, F of [3,2,1,0]:
    #    bytes F[2]16 per group = 96+96; ops end=2, nre=1: F[1]4 per group = 20+20
    quirks=[(928, 212), (832, 216), (912, 208)], quirks_choice=2,
    # corrected: recounted per candidate, memo level b-1-i
    # c0 bytes (F[2]+F[1])16 = 176; c1 F[2]16 = 96, ops 24+24;
    # c2 csf 392+360, ops 92+84, bytes F_C[1]16 + F_D[1]16, ops 20+20
    fresh=[(896, 212), (816, 216), (912, 216)], fresh_choice=0)
KA_N5 = dict(
    dims=[6, 5, 4, 3, 2],
    coords=[(0, 0, 0, 0, 0), (0, 0, 1, 2, 1), (0, 2, 1, 0, 0), (3, 0, 0, 0, 1),
            (3, 0, 0, 1, 1), (5, 4, 3, 2, 0)],
    fibers={(0, 1, 2, 3, 4): [3, 4, 5, 6, 6], (4, 3, 2, 1, 0): [2, 5, 6, 6, 6],
            (0, 1, 2, 4, 3): [3, 4, 5, 5, 6], (3, 4, 2, 1, 0): [3, 5, 6, 6, 6]},
    # c0 [(0,4,3) [0..4], (4,1,0) [4..0]]: csf 416+432, ops 96+100;
    #    bytes (F[4]+F[3]+F[2])16 = 272; ops end=2, nre=3: 16+20+24
    # c1 [(0,4,2),..] stale F=(2,5,6,6,6): bytes (6+6)16; ops end=3: 24+24+24
    # c2 [(0,4,1),..] stale: bytes 6*16; ops end=4: 24*3
    # c3 [(0,3,2) [0,1,2,4,3], (3,2,1) [3,4,2,1,0]] stale: g0 bytes (F[3]+F[2])16
    #    = 192, ops end=2, nre=2: F[1]4 + F[2]4 = 20+24; g1 bytes F[2]16 = 96,
    #    ops end=2, nre=1: F[1]4 = 20
    quirks=[(1120, 256), (1040, 268), (944, 268), (1136, 260)], quirks_choice=0,
    # corrected: c0 bytes (F[3]+F[2]+F[1])16 = 240; c1 (6+5)16, ops 20+20+24;
    # c2 6*16, ops 24*3; c3 csf 400+448, ops 92+104, g0 bytes (5+4)16, ops 16+20,
    # g1 bytes 5*16, ops 20
    fresh=[(1088, 256), (1024, 260), (944, 268), (1072, 252)], fresh_choice=3)


@pytest.mark.parametrize("ka", [KA_N4, KA_N5], ids=["N4", "N5"])
def test_analytic_model_known_answers(ka):
    dims, coords = ka["dims"], ka["coords"]
    N = len(dims)
    inds = np.array(coords, dtype=np.int64).T
    vals = np.arange(1, inds.shape[1] + 1, dtype=np.float64)
    for o, F in ka["fibers"].items():  # the hand counts themselves
        assert O.count_fibers(inds, list(o)) == F, o
    t = A.Tensor.from_arrays(inds, vals, dims)
    for quirks, key in ((True, "quirks"), (False, "fresh")):
        cands, ch = A.plan_analytic(t, 2, ref_quirks=quirks, device=-1)
        assert [(p.pspace, p.pops) for p in cands] == ka[key], key
        assert ch == ka[key + "_choice"], key
        # the Python restatement gives the same numbers
        configs = O.enumerate_plans(N, dims)
        pred = O.predict(configs, lambda o: O.count_fibers(inds, list(o)), inds.shape[1], N,
                         2, quirks=quirks)
        assert pred == ka[key] and O.choose(pred) == ka[key + "_choice"], key
    # the orders the candidates use are the ones counted above
    used = {tuple(g[3]) for c in O.enumerate_plans(N, dims) for g in c}
    assert used == set(ka["fibers"])
