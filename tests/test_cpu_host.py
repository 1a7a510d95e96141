"""Host-side construction (CPU only): canonical form, generators, exports."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import adatm_amd as A
from oracle import npref as O

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("dist,vals", [(A.DIST_UNIFORM, A.VALS_U01), (A.DIST_ZIPF, A.VALS_LOGNORMAL),
                                       (A.DIST_CLUSTER, A.VALS_PM1)])
def test_canonicalize_bit_exact(dist, vals):
    spec = A.make_spec([40, 30, 20, 9, 7], 20000, dist=dist, zipf_s=1.3, nclusters=30,
                       sigma=2.5, values=vals, seed=17)
    t = A.Tensor.synth(spec)
    inds0, vals0 = t.inds().astype(np.int64), t.vals.copy()
    maps = t.canonicalize()
    ui, uv, nd, om = O.canonicalize(inds0, vals0, [40, 30, 20, 9, 7])
    assert t.dims == nd
    assert np.array_equal(t.inds(), ui)
    assert np.array_equal(t.vals, uv)  # duplicates summed in draw order, bit-exact
    for a, b in zip(maps, om):
        assert np.array_equal(a, b)


def test_canonicalize_edge_cases():
    # all nonzeros identical -> one merged entry; a mode of length one
    inds = np.zeros((3, 50), dtype=np.int64)
    inds[1] = 4
    vals = np.arange(50, dtype=np.float64) * 0.1
    t = A.Tensor.from_arrays(inds, vals, [1, 9, 3])
    t.canonicalize()
    assert t.nnz == 1 and t.dims == [1, 1, 1]
    acc = 0.0
    for v in vals:
        acc += v
    assert t.vals[0] == acc
    # out-of-range index is rejected
    bad = A.Tensor.from_arrays(np.array([[0, 5], [0, 0], [0, 0]]), np.ones(2), [3, 2, 2])
    with pytest.raises(A.AdatmError):
        bad.canonicalize()


def test_generator_is_counter_based():
    """A prefix of the draw stream does not depend on the number of draws."""
    a = A.Tensor.synth(A.named_spec(2, 5000))
    b = A.Tensor.synth(A.named_spec(2, 2000))
    assert np.array_equal(a.inds()[:, :2000], b.inds())
    assert np.array_equal(a.vals[:2000], b.vals)


def test_zipf_frequencies():
    spec = A.make_spec([50, 50, 50], 200000, dist=A.DIST_ZIPF, zipf_s=1.1, seed=3)
    t = A.Tensor.synth(spec)
    k = np.arange(1, 51)
    p = k ** -1.1
    p /= p.sum()
    f = np.bincount(t.ind(0), minlength=50) / t.nnz
    assert np.max(np.abs(f - p)) < 0.004


def test_rand_matrix_range_and_determinism():
    a = A.rand_matrix(5, 2, 100, 16)
    b = A.rand_matrix(5, 2, 100, 16)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert not np.array_equal(a, A.rand_matrix(5, 3, 100, 16))


def _declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        txt = re.sub(r"//.*", "", txt)
        txt = re.sub(r"typedef[^;]*;", "", txt)  # function-pointer and struct typedefs
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof", "defined"):
                continue
            names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(A.LIB_PATH))
    declared = _declared_functions()
    assert len(declared) > 40
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_engine_rejects_bad_arguments_before_any_device_call(lib):
    """Shape, rank, tile and plan checks fail with BADINPUT (2) and a message,
    without touching a GPU."""
    import ctypes as C
    spec = A.make_spec([30, 20, 10, 8], 2000, seed=4)
    t = A.Tensor.synth(spec)
    t.canonicalize()
    plan = A.plan_candidates(4, t.dims)[0]
    L = A.lib()
    h = C.c_void_p()
    for rank, tiles in ((65, (0, 0)), (0, (0, 0)), (8, (4, 0)), (8, (0, 5000))):
        rc = L.adatm_engine_create_tiled(C.byref(t.c), rank, C.byref(plan), A.FP32, 0,
                                         tiles[0], tiles[1], C.byref(h))
        assert rc == 2 and not h.value, (rank, tiles, rc)
    bad = A.Plan()
    bad.n_grp = 1
    bad.grp[0].memo_mode, bad.grp[0].begin_imten, bad.grp[0].n_imten = 0, 3, 5
    for l in range(4):
        bad.grp[0].o_opt[l] = l
    rc = L.adatm_engine_create(C.byref(t.c), 8, C.byref(bad), A.FP32, 0, C.byref(h))
    assert rc == 2 and b"stores" in L.adatm_last_error()
    big = A.make_spec([70_000_000, 20, 10, 8], 100, seed=5)
    tb = A.Tensor.synth(big)
    rc = L.adatm_engine_create(C.byref(tb.c), 64, C.byref(plan), A.FP32, 0, C.byref(h))
    assert rc == 2 and b"2^32" in L.adatm_last_error()
    # no nonzeros
    te = A.Tensor.from_arrays(np.zeros((4, 0), dtype=np.uint32), np.zeros(0), [30, 20, 10, 8])
    rc = L.adatm_engine_create(C.byref(te.c), 8, C.byref(plan), A.FP32, 0, C.byref(h))
    assert rc == 2 and not h.value
    # a creation flag this build does not know
    rc = L.adatm_engine_create_ex(C.byref(t.c), 8, C.byref(plan), A.FP32, 0, 0, 0, 0x4,
                                  C.byref(h))
    assert rc == 2 and b"unknown creation flags" in L.adatm_last_error()


def _bf16_rne(u: int) -> int:
    """Nearest-even BF16 of the FP32 with bits u, decided by exact rational
    comparison of the two neighbouring BF16 values (finite inputs)."""
    from fractions import Fraction
    import struct

    def val(bits32):
        return Fraction(struct.unpack("<f", struct.pack("<I", bits32))[0])

    lo = u & 0xFFFF0000
    hi = (lo + 0x10000) & 0xFFFFFFFF
    x = val(u)
    if (hi & 0x7F800000) == 0x7F800000:  # rounding up overflows to infinity
        top = val(lo) + (val(lo) - val(lo - 0x10000))
        d_lo, d_hi = abs(x - val(lo)), abs(top - x)
    else:
        d_lo, d_hi = abs(x - val(lo)), abs(val(hi) - x)
    if d_lo < d_hi or (d_lo == d_hi and ((lo >> 16) & 1) == 0):
        return lo >> 16
    return (hi >> 16) & 0xFFFF


def test_bf16_rounding_bits(lib):
    """The BF16 factor-row rounding: nearest-even on random finite values
    (exact rational reference), infinities kept, every NaN (also one whose
    payload sits in the low 16 mantissa bits only) a quiet NaN of its sign."""
    import struct

    def bits(u):
        return A.bf16_bits(struct.unpack("<f", struct.pack("<I", u))[0])

    rng = np.random.default_rng(5)
    us = list(rng.integers(0, 0x7F7FFFFF, 3000)) + [0x3F808000, 0x3F818000, 0x3F80FFFF,
                                                   0x00000001, 0x7F7FFFFF, 0]
    for u in us:
        for sign in (0, 0x80000000):
            w = int(u) | sign
            assert bits(w) == (_bf16_rne(int(u)) | (sign >> 16)), hex(w)
    assert bits(0x7F800000) == 0x7F80 and bits(0xFF800000) == 0xFF80
    for nan in (0x7F800001, 0x7F80FFFF, 0x7FC00000, 0x7FFFFFFF):
        assert bits(nan) & 0x7FFF > 0x7F80 and bits(nan) & 0x0040 and not bits(nan) & 0x8000
        assert bits(nan | 0x80000000) & 0x8000 and bits(nan | 0x80000000) & 0x7FFF > 0x7F80


def test_autotune_search_order_and_margin(lib):
    """The autotuner's search on a table of timings: the default tiles are
    timed first; a candidate wins only when > 1% faster; when the scatter
    tile moves, the reduction tiles are timed again with it (skipping the one
    already timed); then the root-pass forms, then the batches on the best
    tiles and form."""
    P, W = A.BATCH_PIPELINED, A.BATCH_WIDE
    G0, G2 = A.ENGINE_STAGED_GROUP(0), A.ENGINE_STAGED_GROUP(2)
    table = {(128, 64, 0, 0): 10.0, (32, 64, 0, 0): 9.95,  # within 1%: not taken
             (64, 64, 0, 0): 9.0, (256, 64, 0, 0): 11.0, (512, 64, 0, 0): 12.0,
             (64, 16, 0, 0): 9.5, (64, 32, 0, 0): 8.0, (64, 128, 0, 0): 8.5,
             (64, 256, 0, 0): 9.0, (128, 32, 0, 0): 8.05, (32, 32, 0, 0): 7.0,
             (256, 32, 0, 0): 7.0, (512, 32, 0, 0): 9.0,
             (32, 32, 0, G0 | G2): 6.5, (32, 32, 0, G2): 6.0}
    calls = []

    def timer(r, s, b, f):
        calls.append((r, s, b, f))
        return table.get((r, s, b, f), 6.0 if b == 8 | W else 5.97)

    (r, s, b, f), ms = A.autotune_search(timer, groups=0b101)
    assert calls[:5] == [(128, 64, 0, 0), (32, 64, 0, 0), (64, 64, 0, 0), (256, 64, 0, 0),
                         (512, 64, 0, 0)]
    assert calls[5:9] == [(64, 16, 0, 0), (64, 32, 0, 0), (64, 128, 0, 0), (64, 256, 0, 0)]
    # the scatter tile moved to 32: every reduction tile but 64 again with it
    assert calls[9:13] == [(128, 32, 0, 0), (32, 32, 0, 0), (256, 32, 0, 0), (512, 32, 0, 0)]
    # both memo groups staged (6.5 < 7.0: kept), then group 0 back to one
    # pass (6.0: kept), then group 2 back too (7.0: not kept)
    assert calls[13:16] == [(32, 32, 0, G0 | G2), (32, 32, 0, G2), (32, 32, 0, 0)]
    batches = [4, 8, 16, 4 | P, 8 | P, 4 | W, 8 | W]
    assert calls[16:] == [(32, 32, bt, G2) for bt in batches]
    # 256/32 ties 32/32 (not 1% faster); every batch within 1% of 6.0 loses
    assert (r, s, b, f) == (32, 32, 0, G2) and ms == 6.0
    # the scatter tile stays at its default: no third pass; a batch wins
    calls.clear()
    (r, s, b, f), ms = A.autotune_search(lambda r, s, b, f: calls.append((r, s, b, f)) or
                                         (5.0 if (r, s, b, f) == (128, 64, 8, 0) else 10.0))
    assert len(calls) == 5 + 4 + 7  # no memo group in the mask
    assert (r, s, b, f) == (128, 64, 8, 0) and ms == 5.0


def test_stepwise_cpd_refuses_bad_arguments(lib):
    """cpd_sweep and solve_status check their arguments on the host."""
    import ctypes as C
    L = A.lib()
    fit = C.c_double()
    assert L.adatm_engine_cpd_sweep(None, 0, 0, C.byref(fit)) == A.SPLATT_ERROR_BADINPUT
    assert L.adatm_engine_cpd_begin(None) == A.SPLATT_ERROR_BADINPUT
    st = C.c_int()
    assert L.adatm_engine_solve_status(None, C.byref(st)) == A.SPLATT_ERROR_BADINPUT


def test_persist_plan_balances_chains_and_orders_dependencies():
    """The persistent batched form's queue plan: chains by total cost,
    largest first, onto the least-loaded queue; each queue's stages depth by
    depth with the chains of a depth in assignment order, so a stage always
    follows its chain's previous stage in the same queue."""
    # five one-stage chains of cost 10, 9, 8, 1, 1 on two queues (hand-traced):
    # 10 -> q0, 9 -> q1, 8 -> q1 (9 < 10), 1 -> q0 (10 < 17), 1 -> q0 (11 < 17)
    queue, order, qstart = A.persist_plan([1, 1, 1, 1, 1], [10, 9, 8, 1, 1], 2)
    assert queue.tolist() == [0, 1, 1, 0, 0]
    assert qstart.tolist() == [0, 3, 5]
    assert order.tolist() == [0, 3, 4, 1, 2]
    # chain 0 (3 stages, cost 6) and chain 1 (1 stage, cost 2) on one queue:
    # depth 0 of both (larger chain first), then depths 1 and 2 of chain 0
    queue, order, qstart = A.persist_plan([3, 1], [1, 2, 3, 2], 1)
    assert queue.tolist() == [0, 0] and order.tolist() == [0, 3, 1, 2]
    # empty chains take no stage; equal costs go to the lower queue first
    queue, order, qstart = A.persist_plan([0, 2, 0, 2], [1, 1, 1, 1], 3)
    assert queue.tolist()[1] == 0 and queue.tolist()[3] == 1
    assert qstart.tolist() == [0, 2, 4, 4]
    # random batches: every stage once, chains whole on one queue, each
    # stage after its chain's previous one, loads within one chain of even
    rng = np.random.default_rng(5)
    for _ in range(50):
        n, nq = int(rng.integers(1, 40)), int(rng.integers(1, 9))
        lens = rng.integers(0, 9, n)
        costs = rng.uniform(0.1, 10.0, int(lens.sum()))
        queue, order, qstart = A.persist_plan(lens, costs, nq)
        first = np.concatenate([[0], np.cumsum(lens)[:-1]])
        chain = np.repeat(np.arange(n), lens)
        assert sorted(order.tolist()) == list(range(int(lens.sum())))
        pos = np.empty(len(order), dtype=np.int64)
        pos[order] = np.arange(len(order))
        for q in range(nq):
            for a in order[qstart[q]:qstart[q + 1]]:
                assert queue[chain[a]] == q
                if a > first[chain[a]]:
                    assert pos[a - 1] < pos[a]
        work = np.array([costs[first[i]:first[i] + lens[i]].sum() for i in range(n)])
        load = np.bincount(queue, weights=work, minlength=nq)
        assert load.max() - load.min() <= work.max() + 1e-9
