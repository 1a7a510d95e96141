"""The allocation-free prediction of the cost model against
adatm_mlp_predict, through the C ABI: two calls in a row, the second with
fewer rows, each equal to the allocating entry point's output."""
import ctypes as C
import os

import numpy as np
import pytest

import adatm_amd as A

pytestmark = pytest.mark.gpu

NOALLOC_MAX_ROWS = 4096  # include/adatm.h ADATM_NOALLOC_MAX_ROWS


def _noalloc(model, X):
    L = A.lib()
    f = L.adatm_mlp_predict_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float)]
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.full(X.shape[0], np.nan, dtype=np.float32)
    rc = f(model.h, X.ctypes.data_as(C.POINTER(C.c_float)), X.shape[0],
           y.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, y


def test_predict_noalloc_equals_predict():
    print("no-alloc prediction: trained model, 300 then 37 rows", flush=True)
    path = A.costmodel_path()
    model = A.Mlp(device=0, path=path) if path and os.path.exists(path) else A.Mlp(0, seed=3)
    rng = np.random.default_rng(12)
    xm, xs, _, _ = model.norm()
    for n in (300, 37):
        X = (xm + xs * rng.standard_normal((n, A.MLP_IN))).astype(np.float32)
        ref = model.predict(X)
        rc, y = _noalloc(model, X)
        assert rc == A.SPLATT_SUCCESS, A.lib().adatm_last_error()
        assert np.all(np.isfinite(y))
        np.testing.assert_allclose(y, ref, rtol=1e-6, atol=1e-6)
        assert np.array_equal(y, ref), np.max(np.abs(y - ref))
    big = np.zeros((NOALLOC_MAX_ROWS + 1, A.MLP_IN), dtype=np.float32)
    rc, _ = _noalloc(model, big)
    assert rc == A.SPLATT_ERROR_BADINPUT and b"at most" in A.lib().adatm_last_error()
    model.close()
