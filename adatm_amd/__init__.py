"""MI355X AdaTM: Python mirror of libsplatt_adatm's C API (ctypes).

The compute path is the C library (host C + HIP kernels for gfx950); this
module only marshals arrays.  Loading fails loudly when the library has not
been built -- there is no Python or CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libsplatt_adatm.so"
CLI_PATH = PKG / "bin" / "splatt"

MAX_NMODES = 16
SPLATT_SUCCESS = 1
SPLATT_ERROR_BADINPUT = 2
SPLATT_ERROR_NOMEMORY = 3
SPLATT_ERROR_DEVICE = 4
SPLATT_ERROR_NUMERIC = 5

FP64, FP32, BF16ROWS = 0, 1, 2
ENGINE_STAGED = 0x1
ENGINE_GLOBAL_NODE_ROWS = 0x2  # node rows never through buffer loads


def ENGINE_STAGED_GROUP(g: int) -> int:
    return 0x100 << g
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16rows": BF16ROWS}

DIST_UNIFORM, DIST_ZIPF, DIST_CLUSTER = 0, 1, 2
VALS_U01, VALS_LOGNORMAL, VALS_PM1 = 0, 1, 2

# SPLATT option indices (include/splatt.h)
OPT_NTHREADS, OPT_TOLERANCE, OPT_NITER, OPT_VERBOSITY, OPT_RANDSEED = 0, 1, 2, 3, 4
OPT_CSF_ALLOC, OPT_TILE, OPT_TILEDEPTH, OPT_DECOMP, OPT_COMM = 5, 6, 7, 8, 9
OPT_PRECISION, OPT_PLANNER, OPT_GRAPH, OPT_REFQUIRKS = 10, 11, 12, 13
OPT_SPACE_LIMIT, OPT_DEVICE, OPT_RANK_HINT, OPT_NOPTIONS = 14, 15, 16, 17
PLAN_ANALYTIC, PLAN_LEARNED, PLAN_EXHAUSTIVE = 0, 1, 2
# flag of Engine.set_batch: the next batch's rows load while one folds
BATCH_PIPELINED = 0x100
# flag of Engine.set_batch: 16-byte row pieces per lane (128-bit gathers)
BATCH_WIDE = 0x200
CSF_ONEMODE, CSF_TWOMODE, CSF_ALLMODE = 0, 1, 2
NOTILE, DENSETILE = 0, 1
ADAPTIVE, ADAPTIVE_TIME_EFFICIENT, ADAPTIVE_SPACE_EFFICIENT = 3, 4, 5


class AdatmError(RuntimeError):
    """A libsplatt_adatm entry point returned an error code."""


class Coo(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("nnz", C.c_uint64),
        ("dims", C.c_uint64 * MAX_NMODES),
        ("ind", C.POINTER(C.c_uint32) * MAX_NMODES),
        ("vals", C.POINTER(C.c_double)),
    ]


class SynthSpec(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("dims", C.c_uint64 * MAX_NMODES),
        ("ndraws", C.c_uint64),
        ("index_dist", C.c_int),
        ("zipf_s", C.c_double * MAX_NMODES),
        ("nclusters", C.c_uint64),
        ("cluster_sigma", C.c_double),
        ("value_dist", C.c_int),
        ("seed", C.c_uint64),
    ]


class Group(C.Structure):
    _fields_ = [
        ("memo_mode", C.c_uint32),
        ("begin_imten", C.c_uint32),
        ("n_imten", C.c_uint32),
        ("o_opt", C.c_uint32 * MAX_NMODES),
    ]


class Plan(C.Structure):
    _fields_ = [
        ("n_grp", C.c_uint32),
        ("grp", Group * MAX_NMODES),
        ("pspace", C.c_uint64),
        ("pops", C.c_uint64),
    ]

    def groups(self, nmodes: int) -> list[tuple[int, int, int, list[int]]]:
        return [
            (g.memo_mode, g.begin_imten, g.n_imten, list(g.o_opt[:nmodes]))
            for g in self.grp[: self.n_grp]
        ]


_lib = None


# timer(ctx, reduction tile, scatter tile, batch, flags, &ms) of adatm_autotune_search
TILE_TIMER = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                         C.c_uint32, C.POINTER(C.c_double))


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import "
            "__graft_entry__ as g; g.build()')")
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    sig = {
        "adatm_last_error": (C.c_char_p, []),
        "pair_sort": (None, [P(C.c_uint64), C.c_uint64]),
        "adatm_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "adatm_synth_named": (C.c_int, [C.c_int, C.c_uint64, P(SynthSpec)]),
        "adatm_synth": (C.c_int, [P(SynthSpec), P(Coo)]),
        "adatm_coo_canonicalize": (C.c_int, [P(Coo), P(P(C.c_uint32))]),
        "adatm_coo_free": (None, [P(Coo)]),
        "adatm_coo_alloc": (C.c_int, [C.c_uint32, C.c_uint64, P(C.c_uint64), P(Coo)]),
        "adatm_rand_matrix": (None, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, P(C.c_double)]),
        "adatm_plan_candidates": (C.c_int, [C.c_uint32, P(C.c_uint64), P(Plan), C.c_int]),
        "adatm_count_fibers": (C.c_int, [P(Coo), P(C.c_uint32), P(C.c_uint64), C.c_int]),
        "adatm_plan_analytic": (C.c_int, [P(Coo), C.c_uint64, C.c_int, C.c_double, C.c_int,
                                          P(Plan), C.c_int, P(C.c_int), P(C.c_int)]),
        "adatm_plan_use": (None, [P(Plan), C.c_uint32, P(C.c_uint32), P(C.c_uint32)]),
        "adatm_plan_space_efficient": (C.c_int, [P(Plan), C.c_uint32, P(Plan),
                                                 P(C.c_uint32), P(C.c_uint32)]),
        "adatm_engine_create": (C.c_int, [P(Coo), C.c_uint64, P(Plan), C.c_int, C.c_int,
                                          P(C.c_void_p)]),
        "adatm_engine_destroy": (None, [C.c_void_p]),
        "adatm_engine_create_tiled": (C.c_int, [P(Coo), C.c_uint64, P(Plan), C.c_int, C.c_int,
                                                C.c_uint32, C.c_uint32, P(C.c_void_p)]),
        "adatm_engine_create_ex": (C.c_int, [P(Coo), C.c_uint64, P(Plan), C.c_int, C.c_int,
                                             C.c_uint32, C.c_uint32, C.c_uint32,
                                             P(C.c_void_p)]),
        "adatm_autotune_tiles": (C.c_int, [P(Coo), C.c_uint64, P(Plan), C.c_int, C.c_int,
                                           C.c_int, P(C.c_uint32), P(C.c_uint32),
                                           P(C.c_uint32), P(C.c_uint32), P(C.c_double)]),
        "adatm_engine_set_batch": (C.c_int, [C.c_void_p, C.c_uint32]),
        "adatm_engine_set_factor": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "adatm_engine_get_factor": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "adatm_engine_get_lambda": (C.c_int, [C.c_void_p, P(C.c_double)]),
        "adatm_engine_init_factors": (C.c_int, [C.c_void_p, C.c_uint64]),
        "adatm_engine_mttkrp": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "adatm_engine_mttkrp_seq": (C.c_int, [C.c_void_p]),
        "adatm_engine_cpd": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_int,
                                       P(C.c_double), P(C.c_uint32)]),
        "adatm_engine_time": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                        P(C.c_double), P(C.c_double)]),
        "adatm_engine_run_timed": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_int,
                                             P(C.c_double)]),
        "adatm_engine_get_fits": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "adatm_engine_stats": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_double),
                                         P(C.c_double), P(C.c_uint64)]),
        "adatm_engine_get_csf": (C.c_int, [C.c_void_p, C.c_uint32, P(P(C.c_uint32)),
                                           P(P(C.c_uint32)), P(C.c_double)]),
        "adatm_engine_launches": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int),
                                            P(C.c_double), P(C.c_double)]),
        "adatm_batch_create": (C.c_int, [P(C.c_void_p), C.c_int, P(C.c_void_p)]),
        "adatm_batch_destroy": (None, [C.c_void_p]),
        "adatm_batch_cpd": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double), P(C.c_double)]),
        "adatm_batch_sequential": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "adatm_batch_set_persistent": (C.c_int, [C.c_void_p, C.c_int]),
        "adatm_batch_set_form": (C.c_int, [C.c_void_p, C.c_int]),
        "adatm_batch_form": (C.c_int, [C.c_void_p]),
        "adatm_persist_plan": (C.c_int, [C.c_uint32, P(C.c_uint32), P(C.c_double), C.c_uint32,
                                         P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "adatm_batch_persistent": (C.c_int, [C.c_void_p]),
        "adatm_mlp_nparams": (C.c_size_t, []),
        "adatm_mlp_create": (C.c_int, [C.c_int, C.c_uint64, P(C.c_void_p)]),
        "adatm_mlp_destroy": (None, [C.c_void_p]),
        "adatm_mlp_set_params": (C.c_int, [C.c_void_p, P(C.c_float), C.c_size_t]),
        "adatm_mlp_get_params": (C.c_int, [C.c_void_p, P(C.c_float), C.c_size_t]),
        "adatm_mlp_set_norm": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float),
                                         C.c_float, C.c_float]),
        "adatm_mlp_get_norm": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float),
                                         P(C.c_float), P(C.c_float)]),
        "adatm_mlp_loss_grad": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float),
                                          C.c_uint32, P(C.c_float), P(C.c_float)]),
        "adatm_mlp_step": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_uint32,
                                     C.c_float, P(C.c_float)]),
        "adatm_mlp_train": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                      P(C.c_float)]),
        "adatm_mlp_predict": (C.c_int, [C.c_void_p, P(C.c_float), C.c_uint32,
                                        P(C.c_float)]),
        "adatm_mlp_save": (C.c_int, [C.c_void_p, C.c_char_p]),
        "adatm_mlp_load": (C.c_int, [C.c_int, C.c_char_p, P(C.c_void_p)]),
        "adatm_plan_features": (C.c_int, [P(Coo), C.c_uint64, P(Plan), C.c_int, C.c_int,
                                          P(C.c_float)]),
        "adatm_plan_features_from_counts": (C.c_int, [C.c_uint32, C.c_uint64,
                                                      P(C.c_uint64), C.c_uint64, P(Plan),
                                                      P(C.c_uint64), P(C.c_float)]),
        "adatm_plan_learned": (C.c_int, [P(Coo), C.c_uint64, C.c_void_p, C.c_int,
                                         P(Plan), C.c_int, P(C.c_int), P(C.c_int),
                                         P(C.c_float)]),
        "adatm_plan_exhaustive": (C.c_int, [P(Coo), C.c_uint64, C.c_int, C.c_int, C.c_int,
                                            C.c_int, P(Plan), C.c_int, P(C.c_int),
                                            P(C.c_int), P(C.c_double)]),
        "adatm_costmodel_path": (C.c_int, [C.c_char_p, C.c_size_t]),
        "adatm_engine_cpd_begin": (C.c_int, [C.c_void_p]),
        "adatm_engine_cpd_sweep": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int,
                                             P(C.c_double)]),
        "adatm_engine_solve_status": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "adatm_autotune_search": (C.c_int, [TILE_TIMER, C.c_void_p, C.c_uint32, P(C.c_uint32),
                                            P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                            P(C.c_double)]),
        "adatm_k_bf16_bits": (C.c_uint16, [C.c_float]),
        "adatm_perm_nvtxs": (C.c_int, [P(Coo), C.c_int, C.c_uint32, C.c_int,
                                       P(C.c_uint64)]),
        "adatm_perm_compute": (C.c_int, [P(Coo), C.c_int, C.c_uint32, P(C.c_uint32),
                                         C.c_uint64, C.c_int, P(P(C.c_uint32))]),
        "adatm_coo_relabel": (None, [P(Coo), P(P(C.c_uint32))]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != SPLATT_SUCCESS:
        msg = lib().adatm_last_error().decode(errors="replace")
        raise AdatmError(f"{what} failed with code {rc}: {msg}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _u64ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


# --------------------------------------------------------------------------
class Tensor:
    """A coordinate tensor owned by the C library (u32 indices, f64 values)."""

    def __init__(self, coo: Coo):
        self.c = coo

    @classmethod
    def from_arrays(cls, inds: np.ndarray, vals: np.ndarray, dims) -> "Tensor":
        inds = np.asarray(inds)
        N, nnz = inds.shape
        coo = Coo()
        d = np.asarray(dims, dtype=np.uint64)
        check(lib().adatm_coo_alloc(N, nnz, _u64ptr(d), C.byref(coo)), "coo_alloc")
        t = cls(coo)
        for m in range(N):
            t.ind(m)[:] = inds[m]
        t.vals[:] = vals
        return t

    @classmethod
    def synth(cls, spec: SynthSpec) -> "Tensor":
        coo = Coo()
        check(lib().adatm_synth(C.byref(spec), C.byref(coo)), "synth")
        return cls(coo)

    @property
    def nmodes(self) -> int:
        return int(self.c.nmodes)

    @property
    def nnz(self) -> int:
        return int(self.c.nnz)

    @property
    def dims(self) -> list[int]:
        return [int(x) for x in self.c.dims[: self.nmodes]]

    def ind(self, m: int) -> np.ndarray:
        return np.ctypeslib.as_array(self.c.ind[m], shape=(self.nnz,))

    @property
    def vals(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.c.vals, shape=(self.nnz,))

    def inds(self) -> np.ndarray:
        return np.stack([self.ind(m) for m in range(self.nmodes)])

    def canonicalize(self) -> list[np.ndarray]:
        """Merge duplicates, compact empty indices; returns indmap per mode."""
        maps = (C.POINTER(C.c_uint32) * MAX_NMODES)()
        old_dims = self.dims
        check(lib().adatm_coo_canonicalize(C.byref(self.c), maps), "canonicalize")
        out = []
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for m in range(self.nmodes):
            n = self.dims[m]
            arr = np.ctypeslib.as_array(maps[m], shape=(max(n, 1),))[:n].copy()
            libc.free(C.cast(maps[m], C.c_void_p))
            out.append(arr)
        self.orig_dims = old_dims
        return out

    def free(self) -> None:
        if self.c is not None and self.c.nmodes:
            lib().adatm_coo_free(C.byref(self.c))
            self.c.nmodes = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def named_spec(config: int, ndraws: int = 0) -> SynthSpec:
    spec = SynthSpec()
    check(lib().adatm_synth_named(config, ndraws, C.byref(spec)), "synth_named")
    return spec


def make_spec(dims, ndraws, dist=DIST_UNIFORM, zipf_s=1.1, nclusters=0, sigma=0.0,
              values=VALS_U01, seed=1) -> SynthSpec:
    spec = SynthSpec()
    spec.nmodes = len(dims)
    for m, d in enumerate(dims):
        spec.dims[m] = int(d)
        spec.zipf_s[m] = float(zipf_s[m] if np.ndim(zipf_s) else zipf_s)
    spec.ndraws = int(ndraws)
    spec.index_dist = dist
    spec.nclusters = int(nclusters)
    spec.cluster_sigma = float(sigma)
    spec.value_dist = values
    spec.seed = int(seed)
    return spec


def synth_tensor(config: int, ndraws: int = 0, canonical: bool = True) -> Tensor:
    t = Tensor.synth(named_spec(config, ndraws))
    if canonical:
        t.canonicalize()
    return t


def rand_matrix(seed: int, mode: int, rows: int, rank: int) -> np.ndarray:
    out = np.empty((rows, rank), dtype=np.float64)
    lib().adatm_rand_matrix(seed, mode, rows, rank, _dptr(out))
    return out


def philox(ctr, key) -> list[int]:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().adatm_philox4x32_10(_uptr(c), _uptr(k), _uptr(o))
    return [int(x) for x in o]


# --------------------------------------------------------------------------
def plan_candidates(nmodes: int, dims) -> list[Plan]:
    d = np.asarray(list(dims) + [0] * (MAX_NMODES - len(dims)), dtype=np.uint64)
    arr = (Plan * 256)()
    n = lib().adatm_plan_candidates(nmodes, _u64ptr(d), arr, 256)
    return [arr[i] for i in range(n)]


PERM_GRAPH, PERM_HGRAPH, PERM_FIBSCHED, PERM_ERROR = 0, 1, 2, 3


def perm_nvtxs(t: Tensor, kind: int, mode: int = 0, device: int = -1) -> int:
    """Vertices of the partition that relabelling `kind` reads."""
    n = C.c_uint64(0)
    check(lib().adatm_perm_nvtxs(C.byref(t.c), kind, mode, device, C.byref(n)),
          "perm_nvtxs")
    return int(n.value)


def perm_compute(t: Tensor, kind: int, parts, mode: int = 0,
                 device: int = -1) -> list[np.ndarray]:
    """New label of every index (old -> new), one array per mode."""
    parts = np.ascontiguousarray(parts, dtype=np.uint32)
    out = [np.zeros(max(d, 1), dtype=np.uint32) for d in t.dims]
    ptrs = (C.POINTER(C.c_uint32) * t.nmodes)(*[_uptr(a) for a in out])
    check(lib().adatm_perm_compute(C.byref(t.c), kind, mode, _uptr(parts), parts.size,
                                   device, ptrs), "perm_compute")
    return [a[:d] for a, d in zip(out, t.dims)]


def relabel(t: Tensor, perms) -> None:
    """t.ind[m][j] = perms[m][t.ind[m][j]] (None leaves a mode alone)."""
    keep = [None if p is None else np.ascontiguousarray(p, dtype=np.uint32) for p in perms]
    ptrs = (C.POINTER(C.c_uint32) * t.nmodes)(
        *[C.POINTER(C.c_uint32)() if p is None else _uptr(p) for p in keep])
    lib().adatm_coo_relabel(C.byref(t.c), ptrs)


def count_fibers(t: Tensor, order, device: int = -1) -> list[int]:
    o = np.asarray(list(order), dtype=np.uint32)
    F = np.zeros(MAX_NMODES, dtype=np.uint64)
    check(lib().adatm_count_fibers(C.byref(t.c), _uptr(o), _u64ptr(F), device),
          "count_fibers")
    return [int(x) for x in F[: t.nmodes]]


def plan_analytic(t: Tensor, rank: int, ref_quirks: bool = True,
                  space_limit: float = 90e9, device: int = -1):
    """(candidates, chosen index) of the analytic planner."""
    arr = (Plan * 256)()
    nc = C.c_int()
    ch = C.c_int()
    check(lib().adatm_plan_analytic(C.byref(t.c), rank, int(ref_quirks), space_limit,
                                    device, arr, 256, C.byref(nc), C.byref(ch)),
          "plan_analytic")
    return [arr[i] for i in range(nc.value)], ch.value


def plan_space_efficient(plan: Plan, nmodes: int):
    """(kept plan, use groups, use tags) of SPACE_EFFICIENT allocation."""
    out = Plan()
    g = np.zeros(MAX_NMODES, dtype=np.uint32)
    t = np.zeros(MAX_NMODES, dtype=np.uint32)
    lib().adatm_plan_space_efficient(C.byref(plan), nmodes, C.byref(out), _uptr(g), _uptr(t))
    return out, [int(x) for x in g[:nmodes]], [int(x) for x in t[:nmodes]]


def plan_use(plan: Plan, nmodes: int):
    g = np.zeros(MAX_NMODES, dtype=np.uint32)
    t = np.zeros(MAX_NMODES, dtype=np.uint32)
    lib().adatm_plan_use(C.byref(plan), nmodes, _uptr(g), _uptr(t))
    return [int(x) for x in g[:nmodes]], [int(x) for x in t[:nmodes]]


MLP_IN = 24


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def plan_features(t: Tensor, rank: int, plans, device: int = -1) -> np.ndarray:
    """(len(plans), 24) features of the learned cost model."""
    arr = (Plan * len(plans))(*plans)
    f = np.zeros((len(plans), MLP_IN), dtype=np.float32)
    check(lib().adatm_plan_features(C.byref(t.c), rank, arr, len(plans), device,
                                    _fptr(f)), "plan_features")
    return f


def plan_features_from_counts(nmodes: int, nnz: int, dims, rank: int, plan: Plan,
                              Fg) -> np.ndarray:
    d = np.asarray(list(dims), dtype=np.uint64)
    F = np.ascontiguousarray(np.asarray(Fg, dtype=np.uint64).reshape(-1))
    f = np.zeros(MLP_IN, dtype=np.float32)
    check(lib().adatm_plan_features_from_counts(nmodes, nnz, _u64ptr(d), rank,
                                                C.byref(plan), _u64ptr(F), _fptr(f)),
          "plan_features_from_counts")
    return f


class Mlp:
    """The learned cost model (wraps adatm_mlp; weights on the GPU)."""

    def __init__(self, device: int = 0, seed: int = 1, path=None):
        self.h = C.c_void_p()
        if path is not None:
            check(lib().adatm_mlp_load(device, str(path).encode(), C.byref(self.h)),
                  "mlp_load")
        else:
            check(lib().adatm_mlp_create(device, seed, C.byref(self.h)), "mlp_create")

    @staticmethod
    def nparams() -> int:
        return int(lib().adatm_mlp_nparams())

    def close(self):
        if self.h:
            lib().adatm_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams(), dtype=np.float32)
        check(lib().adatm_mlp_get_params(self.h, _fptr(p), p.size), "mlp_get_params")
        return p

    def set_params(self, p) -> None:
        p = _f32(p)
        check(lib().adatm_mlp_set_params(self.h, _fptr(p), p.size), "mlp_set_params")

    def set_norm(self, xmean, xstd, ymean: float, ystd: float) -> None:
        a, b = _f32(xmean), _f32(xstd)
        check(lib().adatm_mlp_set_norm(self.h, _fptr(a), _fptr(b), ymean, ystd),
              "mlp_set_norm")

    def norm(self):
        a = np.zeros(MLP_IN, dtype=np.float32)
        b = np.zeros(MLP_IN, dtype=np.float32)
        ym, ys = C.c_float(), C.c_float()
        check(lib().adatm_mlp_get_norm(self.h, _fptr(a), _fptr(b), C.byref(ym),
                                       C.byref(ys)), "mlp_get_norm")
        return a, b, ym.value, ys.value

    def loss_grad(self, X, y):
        X, y = _f32(X), _f32(y)
        g = np.zeros(self.nparams(), dtype=np.float32)
        loss = C.c_float()
        check(lib().adatm_mlp_loss_grad(self.h, _fptr(X), _fptr(y), len(y), _fptr(g),
                                        C.byref(loss)), "mlp_loss_grad")
        return loss.value, g

    def step(self, X, y, lr: float) -> float:
        X, y = _f32(X), _f32(y)
        loss = C.c_float()
        check(lib().adatm_mlp_step(self.h, _fptr(X), _fptr(y), len(y), lr,
                                   C.byref(loss)), "mlp_step")
        return loss.value

    def train(self, X, y, epochs: int, batch: int = 64, lr: float = 1e-3,
              seed: int = 1) -> np.ndarray:
        X, y = _f32(X), _f32(y)
        hist = np.zeros(epochs, dtype=np.float32)
        check(lib().adatm_mlp_train(self.h, _fptr(X), _fptr(y), len(y), epochs, batch,
                                    lr, seed, _fptr(hist)), "mlp_train")
        return hist

    def predict(self, X) -> np.ndarray:
        X = _f32(X)
        out = np.zeros(X.shape[0], dtype=np.float32)
        check(lib().adatm_mlp_predict(self.h, _fptr(X), X.shape[0], _fptr(out)),
              "mlp_predict")
        return out

    def save(self, path) -> None:
        check(lib().adatm_mlp_save(self.h, str(path).encode()), "mlp_save")


def plan_learned(t: Tensor, rank: int, model: "Mlp | None", device: int = 0):
    """(candidates, chosen index, predictions) of the learned planner; the
    analytic choice when model is None."""
    arr = (Plan * 256)()
    nc = C.c_int()
    ch = C.c_int()
    pred = np.zeros(256, dtype=np.float32)
    check(lib().adatm_plan_learned(C.byref(t.c), rank, model.h if model else None,
                                   device, arr, 256, C.byref(nc), C.byref(ch),
                                   _fptr(pred)), "plan_learned")
    return [arr[i] for i in range(nc.value)], ch.value, pred[: nc.value]


def plan_exhaustive(t: Tensor, rank: int, precision: int = FP32, device: int = 0,
                    nwarm: int = 2, nrep: int = 5):
    """(candidates, fastest index, ms per candidate) timing one CP-ALS sweep
    (hipGraph) of every candidate on the device."""
    arr = (Plan * 256)()
    nc = C.c_int()
    ch = C.c_int()
    ms = np.zeros(256, dtype=np.float64)
    check(lib().adatm_plan_exhaustive(C.byref(t.c), rank, precision, device, nwarm, nrep,
                                      arr, 256, C.byref(nc), C.byref(ch), _dptr(ms)),
          "plan_exhaustive")
    return [arr[i] for i in range(nc.value)], ch.value, ms[: nc.value].copy()


def autotune_tiles(t: Tensor, rank: int, plan: Plan, precision: int = FP32,
                   device: int = 0, nrep: int = 3):
    """((reduction tile, scatter tile, children per batch, creation flags),
    best ms per sweep) by timing; pass the tuple as Engine(..., tiles=...)."""
    r, s_, b, f = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    ms = C.c_double()
    check(lib().adatm_autotune_tiles(C.byref(t.c), rank, C.byref(plan), precision, device,
                                     nrep, C.byref(r), C.byref(s_), C.byref(b), C.byref(f),
                                     C.byref(ms)),
          "autotune_tiles")
    return (r.value, s_.value, b.value, f.value), ms.value


def autotune_search(timer, groups: int = 0):
    """The autotuner's search driven by timer(red_tile, scat_tile, batch,
    flags) -> ms over the memo groups in the mask `groups`; returns
    ((red_tile, scat_tile, batch, flags), best ms)."""
    def cb(_ctx, r, s_, b, f, ms):
        ms[0] = float(timer(int(r), int(s_), int(b), int(f)))
        return SPLATT_SUCCESS
    fn = TILE_TIMER(cb)
    r, s_, b, f, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
    check(lib().adatm_autotune_search(fn, None, groups, C.byref(r), C.byref(s_), C.byref(b),
                                      C.byref(f), C.byref(ms)), "autotune_search")
    return (r.value, s_.value, b.value, f.value), ms.value


def bf16_bits(x: float) -> int:
    """BF16 bits of an FP32 value as the BF16 factor-row copies round it"""
    return int(lib().adatm_k_bf16_bits(C.c_float(x)))


def costmodel_path():
    buf = C.create_string_buffer(4096)
    return buf.value.decode() if lib().adatm_costmodel_path(buf, 4096) else None


# --------------------------------------------------------------------------
class Engine:
    """A memoization plan resident on one GPU (wraps adatm_engine)."""

    def __init__(self, t: Tensor, rank: int, plan: Plan, precision: int = FP64,
                 device: int = 0, tiles=(0, 0), staged: bool = False, flags: int = 0):
        """staged: root MTTKRPs stop at every memo level (ENGINE_STAGED)
        instead of one leaf-to-root pass writing the memo levels beside;
        flags: any other ENGINE_* creation flags"""
        self.h = C.c_void_p()
        self.N = t.nmodes
        self.R = int(rank)
        self.dims = t.dims
        self.plan = plan
        self.precision = precision
        if len(tiles) > 3:  # the autotuner's creation flags
            flags |= int(tiles[3])
        check(lib().adatm_engine_create_ex(C.byref(t.c), rank, C.byref(plan), precision,
                                           device, int(tiles[0]), int(tiles[1]),
                                           (ENGINE_STAGED if staged else 0) | flags,
                                           C.byref(self.h)), "engine_create")
        if len(tiles) > 2 and tiles[2]:
            self.set_batch(int(tiles[2]))

    def set_batch(self, batch: int) -> None:
        """children per reduction batch: 0 (default), 4, 8 or 16, optionally
        | BATCH_PIPELINED or | BATCH_WIDE"""
        check(lib().adatm_engine_set_batch(self.h, batch), "engine_set_batch")

    def close(self) -> None:
        if self.h:
            lib().adatm_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_factor(self, m: int, A: np.ndarray) -> None:
        A = np.ascontiguousarray(A, dtype=np.float64)
        assert A.shape == (self.dims[m], self.R)
        check(lib().adatm_engine_set_factor(self.h, m, _dptr(A)), "set_factor")

    def get_factor(self, m: int) -> np.ndarray:
        A = np.empty((self.dims[m], self.R), dtype=np.float64)
        check(lib().adatm_engine_get_factor(self.h, m, _dptr(A)), "get_factor")
        return A

    def get_lambda(self) -> np.ndarray:
        lam = np.empty(self.R, dtype=np.float64)
        check(lib().adatm_engine_get_lambda(self.h, _dptr(lam)), "get_lambda")
        return lam

    def init_factors(self, seed: int) -> None:
        check(lib().adatm_engine_init_factors(self.h, seed), "init_factors")

    def mttkrp(self, m: int, out: bool = True):
        M = np.empty((self.dims[m], self.R), dtype=np.float64) if out else None
        check(lib().adatm_engine_mttkrp(self.h, m, _dptr(M) if out else None), "mttkrp")
        return M

    def mttkrp_seq(self) -> None:
        check(lib().adatm_engine_mttkrp_seq(self.h), "mttkrp_seq")

    def cpd(self, niters: int, tol: float = 0.0, graph: bool = False):
        fits = np.zeros(niters, dtype=np.float64)
        done = C.c_uint32()
        check(lib().adatm_engine_cpd(self.h, niters, tol, int(graph), _dptr(fits),
                                     C.byref(done)), "cpd")
        return fits[: done.value]

    def cpd_begin(self) -> None:
        check(lib().adatm_engine_cpd_begin(self.h), "cpd_begin")

    def cpd_sweep(self, it: int, graph: bool = False) -> float:
        fit = C.c_double()
        check(lib().adatm_engine_cpd_sweep(self.h, it, int(graph), C.byref(fit)),
              "cpd_sweep")
        return fit.value

    def solve_status(self) -> int:
        """worst (hadamard of Grams)^-1 status since the last CP-ALS start:
        0 clean, 1 solved with a ridge, 2 failed"""
        st = C.c_int()
        check(lib().adatm_engine_solve_status(self.h, C.byref(st)), "solve_status")
        return st.value

    def time(self, what: int, nwarm: int, nrep: int, graph: bool = False,
             per_kernel: bool = True):
        """Median ms of `what` (0 MTTKRP sequence, 1 ALS sweep); with
        per_kernel, also the per-class totals of one instrumented pass."""
        ms = C.c_double()
        pk = np.zeros(24, dtype=np.float64)
        check(lib().adatm_engine_time(self.h, what, nwarm, nrep, int(graph),
                                      C.byref(ms), _dptr(pk) if per_kernel else None),
              "engine_time")
        return ms.value, pk

    def launches(self, max_n: int = 1024):
        """(class, info, ms, algorithmic bytes) of every launch of the
        instrumented pass of the last time() call; info = levels fused by a
        reduction or the level of a scatter."""
        cls = np.zeros(max_n, dtype=np.int32)
        info = np.zeros(max_n, dtype=np.int32)
        ms = np.zeros(max_n, dtype=np.float64)
        by = np.zeros(max_n, dtype=np.float64)
        ip = C.POINTER(C.c_int)
        n = lib().adatm_engine_launches(self.h, max_n, cls.ctypes.data_as(ip),
                                        info.ctypes.data_as(ip), _dptr(ms), _dptr(by))
        return [(int(cls[i]), int(info[i]), float(ms[i]), float(by[i]))
                for i in range(max(n, 0))]

    def run_timed(self, what: int, nsteps: int, graph: bool = False) -> float:
        ms = C.c_double()
        check(lib().adatm_engine_run_timed(self.h, what, nsteps, int(graph), C.byref(ms)),
              "run_timed")
        return ms.value

    def fits(self, n: int) -> np.ndarray:
        f = np.zeros(n, dtype=np.float64)
        check(lib().adatm_engine_get_fits(self.h, n, _dptr(f)), "get_fits")
        return f

    def stats(self):
        F = np.zeros(self.plan.n_grp * self.N, dtype=np.uint64)
        b = C.c_double()
        f = C.c_double()
        db = C.c_uint64()
        check(lib().adatm_engine_stats(self.h, _u64ptr(F), C.byref(b), C.byref(f),
                                       C.byref(db)), "engine_stats")
        return F.reshape(self.plan.n_grp, self.N), b.value, f.value, db.value

    def get_csf(self, g: int, nnz: int):
        F, _, _, _ = self.stats()
        F = [int(x) for x in F[g]]
        N = self.N
        fptr = [np.zeros(F[l] + 1, dtype=np.uint32) if l < N - 1 else None
                for l in range(N)]
        fids = [np.zeros(F[l], dtype=np.uint32) if l >= 1 else None for l in range(N)]
        pf = (C.POINTER(C.c_uint32) * MAX_NMODES)()
        pi = (C.POINTER(C.c_uint32) * MAX_NMODES)()
        for l in range(N):
            if fptr[l] is not None:
                pf[l] = _uptr(fptr[l])
            if fids[l] is not None:
                pi[l] = _uptr(fids[l])
        vals = np.zeros(nnz, dtype=np.float64)
        check(lib().adatm_engine_get_csf(self.h, g, pf, pi, _dptr(vals)), "get_csf")
        return F, fptr, fids, vals


def persist_plan(lens, costs, nq: int):
    """Queue plan of one round of the persistent batched form
    (adatm_persist_plan): chains of lens[i] stages with costs stored chain
    after chain -> (queue of each chain, stage order, start of each queue)."""
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    costs = np.ascontiguousarray(costs, dtype=np.float64)
    n = len(lens)
    queue = np.zeros(max(n, 1), dtype=np.uint32)
    order = np.zeros(max(int(lens.sum()), 1), dtype=np.uint32)
    qstart = np.zeros(nq + 1, dtype=np.uint32)
    u32 = C.POINTER(C.c_uint32)
    check(lib().adatm_persist_plan(n, lens.ctypes.data_as(u32), _dptr(costs), nq,
                                   queue.ctypes.data_as(u32), order.ctypes.data_as(u32),
                                   qstart.ctypes.data_as(u32)), "persist_plan")
    return queue[:n], order[: int(lens.sum())], qstart


# forms of a batched sweep's MTTKRPs (adatm_batch_set_form)
FORM_FORKED, FORM_PERSISTENT, FORM_GROUPED = 0, 1, 2


class Batch:
    """Batched CP-ALS over independent engines (wraps adatm_batch); keeps
    the engines alive while it exists."""

    def __init__(self, engines):
        self.engines = list(engines)
        self.h = C.c_void_p()
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        check(lib().adatm_batch_create(arr, len(self.engines), C.byref(self.h)),
              "batch_create")

    def cpd(self, niters: int):
        """(device ms of niters batched sweeps, fits (n, niters))."""
        ms = C.c_double()
        fits = np.zeros((len(self.engines), niters), dtype=np.float64)
        check(lib().adatm_batch_cpd(self.h, niters, C.byref(ms), _dptr(fits)), "batch_cpd")
        return ms.value, fits

    def set_persistent(self, on: bool) -> None:
        """one persistent launch per round (True) or one forked branch of
        launches per instance (False)"""
        check(lib().adatm_batch_set_persistent(self.h, int(bool(on))), "batch_set_persistent")

    @property
    def persistent(self) -> bool:
        """whether the last batched sweep ran in persistent mode"""
        return bool(lib().adatm_batch_persistent(self.h))

    def set_form(self, form: int) -> None:
        """FORM_GROUPED (default where it applies), FORM_PERSISTENT or
        FORM_FORKED"""
        check(lib().adatm_batch_set_form(self.h, int(form)), "batch_set_form")

    @property
    def form(self) -> int:
        """the form of the last built batched sweep"""
        return int(lib().adatm_batch_form(self.h))

    def sequential(self, niters: int) -> float:
        ms = C.c_double()
        check(lib().adatm_batch_sequential(self.h, niters, C.byref(ms)), "batch_sequential")
        return ms.value

    def close(self):
        if self.h:
            lib().adatm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
