"""vt_numerical_flux_noalloc (the Roe flux staged through the library's static
device array, no allocation) against vt_numerical_flux (per-call device
buffers) through the C ABI: bit-identical fluxes and equal error codes on two
calls in a row, the second smaller than the first, and sizes above the
array's limit refused."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_RECORDS = 32768


def _records(seed, n):
    """Roe-flux inputs UL[4] UR[4] n[2]: states near the vortex's range, a
    few with a non-physical density or pressure so that codes 1-3 occur."""
    g = np.random.default_rng(seed)
    rec = np.zeros((n, 10))
    for side in (0, 4):
        rho = g.uniform(0.6, 1.4, n)
        u, v = g.uniform(-1.0, 1.5, n), g.uniform(-1.0, 1.5, n)
        p = g.uniform(1.0, 4.0, n)
        rec[:, side] = rho
        rec[:, side + 1] = rho * u
        rec[:, side + 2] = rho * v
        rec[:, side + 3] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    th = g.uniform(0, 2 * np.pi, n)
    rec[:, 8], rec[:, 9] = np.cos(th), np.sin(th)
    bad = g.choice(n, size=max(1, n // 50), replace=False)
    rec[bad[0::3], 0] *= -1.0          # left density negative
    rec[bad[1::3], 7] = 0.01           # right pressure negative
    rec[bad[2::3], 3] = 0.5 * rec[bad[2::3], 3]
    return np.ascontiguousarray(rec)


def _lib(vt):
    L = vt.lib()
    f = L.vt_numerical_flux_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                  ctypes.c_int]
    return L


def _call(fn, rec):
    n = rec.shape[0]
    flux = np.full((n, 4), np.nan)
    codes = np.full(n, -7, dtype=np.int32)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    rc = fn(rec.ctypes.data_as(dp), flux.ctypes.data_as(dp), codes.ctypes.data_as(ip), n)
    return rc, flux, codes


def test_noalloc_flux_equals_allocating_flux_on_two_calls(vt, gpu):
    L = _lib(vt)
    for seed, n in ((1, MAX_RECORDS), (2, 5003)):
        print("no-alloc flux: seed %d, %d records" % (seed, n), flush=True)
        rec = _records(seed, n)
        rc0, f0, c0 = _call(L.vt_numerical_flux, rec)
        rc1, f1, c1 = _call(L.vt_numerical_flux_noalloc, rec)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(c0, c1)
        assert set(np.unique(c1)) >= {0, 1}
        assert np.array_equal(f0, f1, equal_nan=True)
        assert np.isfinite(f1[c1 == 0]).all()


def test_noalloc_flux_on_reference_known_answers(vt, gpu):
    """The reference's known answers (whose agreement with the allocating
    entry point is checked in test_gpu_parity): bit-identical to it."""
    L = _lib(vt)
    d = np.load(os.path.join(HERE, "golden", "roe_flux_known_answers.npz"))
    rec = np.ascontiguousarray(d["records"])
    rc0, f0, c0 = _call(L.vt_numerical_flux, rec)
    rc1, f1, c1 = _call(L.vt_numerical_flux_noalloc, rec)
    assert rc0 == 0 and rc1 == 0 and (c1 == 0).all()
    assert np.array_equal(f0, f1)
    assert np.median(np.abs(f1 - d["flux"]) / np.maximum(np.abs(d["flux"]), 1.0)) < 1e-15


def test_noalloc_flux_refuses_sizes_above_the_limit(vt, gpu):
    L = _lib(vt)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    one = np.zeros(10)
    out = np.zeros(4)
    code = np.zeros(1, dtype=np.int32)
    args = (one.ctypes.data_as(dp), out.ctypes.data_as(dp), code.ctypes.data_as(ip))
    assert L.vt_numerical_flux_noalloc(*args, MAX_RECORDS + 1) == vt.VT_ERR_ARG
    assert L.vt_numerical_flux_noalloc(*args, -1) == vt.VT_ERR_ARG
    assert L.vt_numerical_flux_noalloc(None, None, None, 0) == 0
