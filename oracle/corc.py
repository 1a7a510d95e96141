"""ctypes loader of the oracle's C half (oracle/_build/liboracle.so) --
TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "_build" / "liboracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(SO))
        P = C.POINTER
        L.orc_build.restype = C.c_void_p
        L.orc_build.argtypes = [C.c_uint32, C.c_uint64, P(P(C.c_uint32)), P(C.c_uint64),
                                P(C.c_double), C.c_uint32, P(C.c_uint32), C.c_uint32]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_mttkrp.argtypes = [C.c_void_p, C.c_uint32, P(P(C.c_double)), P(C.c_double)]
        L.orc_cpd.argtypes = [C.c_void_p, P(P(C.c_double)), C.c_uint32, C.c_double,
                              P(C.c_double), P(C.c_double)]
        L.orc_time.restype = C.c_double
        L.orc_time.argtypes = [C.c_void_p, P(P(C.c_double)), C.c_int, C.c_int, C.c_int]
        L.orc_get_nfibs.argtypes = [C.c_void_p, C.c_uint32, P(C.c_uint64)]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


class OracleSeq:
    """FP64 C/OpenMP restatement of the reference CPU path on one plan.

    groups: list of (memo_mode, begin_imten, n_imten, order)."""

    def __init__(self, inds: np.ndarray, vals: np.ndarray, dims, groups, R: int):
        L = lib()
        self.N = inds.shape[0]
        self.R = R
        self.dims = [int(d) for d in dims]
        self._ind = [np.ascontiguousarray(inds[m], dtype=np.uint32) for m in range(self.N)]
        ip = (C.POINTER(C.c_uint32) * self.N)(
            *[a.ctypes.data_as(C.POINTER(C.c_uint32)) for a in self._ind])
        d = np.asarray(self.dims, dtype=np.uint64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        desc = []
        for (mm, b, nim, o) in groups:
            desc += [mm, b, nim] + list(o)
        desc = np.asarray(desc, dtype=np.uint32)
        self.h = L.orc_build(self.N, v.shape[0], ip, d.ctypes.data_as(C.POINTER(C.c_uint64)),
                             v.ctypes.data_as(C.POINTER(C.c_double)), len(groups),
                             desc.ctypes.data_as(C.POINTER(C.c_uint32)), R)
        if not self.h:
            raise ValueError("oracle: bad plan or tensor")
        self.ngrp = len(groups)

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fp(self, U):
        self._keep = [np.array(u, dtype=np.float64, order="C", copy=True) for u in U]
        return (C.POINTER(C.c_double) * self.N)(
            *[u.ctypes.data_as(C.POINTER(C.c_double)) for u in self._keep])

    def nfibs(self, g: int) -> list[int]:
        F = np.zeros(16, dtype=np.uint64)
        lib().orc_get_nfibs(self.h, g, F.ctypes.data_as(C.POINTER(C.c_uint64)))
        return [int(x) for x in F[: self.N]]

    def mttkrp(self, mode: int, U) -> np.ndarray:
        out = np.zeros((self.dims[mode], self.R))
        lib().orc_mttkrp(self.h, mode, self._fp(U), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def cpd(self, U, niters: int, tol: float = 0.0):
        fp = self._fp(U)
        fits = np.zeros(niters)
        lam = np.zeros(self.R)
        n = lib().orc_cpd(self.h, fp, niters, tol, fits.ctypes.data_as(C.POINTER(C.c_double)),
                          lam.ctypes.data_as(C.POINTER(C.c_double)))
        return fits[:n], [u.copy() for u in self._keep], lam

    def time(self, U, what: int, nrep: int, threads: int = 0) -> float:
        return lib().orc_time(self.h, self._fp(U), what, nrep, threads)


def max_threads() -> int:
    return lib().orc_max_threads()
