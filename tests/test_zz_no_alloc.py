"""sbz_solve_batch_noalloc (device scratch array, no allocation) against sbz_solve_batch (plan
buffers from hipMalloc) through the C ABI: two calls in a row, the second smaller, with both
wavefront kernels; final states, saved states, return codes and step statistics must be identical
(same kernels on the same inputs).  Oversized calls fail with an argument error."""
import ctypes

import numpy as np
import pytest

from conftest import Setup
from symboltz import _lib, models
from symboltz.solver import GpuModel, _Opts, ALGO_WAVE, ALGO_WAVE2

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def batch(gm, name, iset, ks, t0, saveat, algo):
    f = getattr(gm.L, name)
    f.argtypes = [ctypes.c_void_p, _ip, _dp, _dp, ctypes.c_int, ctypes.POINTER(_Opts), _dp, ctypes.c_int, _dp, _dp,
                  _ip, _llp]
    n, ns = len(ks), len(saveat)
    u = np.full((n, gm.n), -1.0)
    us = np.full((n, max(ns, 1), gm.n), -1.0)
    ret = np.full(n, -7, dtype=np.int32)
    st = np.zeros((n, 6), dtype=np.int64)
    opts = _Opts(1e-5, 1e-5, 100000, 0, 1 if ns else 0, algo)
    iset = np.ascontiguousarray(iset, dtype=np.int32)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    t0 = np.ascontiguousarray(t0, dtype=np.float64)
    sv = np.ascontiguousarray(saveat if ns else np.zeros(1), dtype=np.float64)
    rc = f(gm.h, _lib.ptr(iset, ctypes.c_int), _lib.ptr(ks), _lib.ptr(t0), n, ctypes.byref(opts), _lib.ptr(sv), ns,
           _lib.ptr(u), _lib.ptr(us) if ns else None, _lib.ptr(ret, ctypes.c_int), _lib.ptr(st, ctypes.c_longlong))
    return rc, u, (us if ns else None), ret, st


def test_noalloc_batch_equals_allocating_batch():
    M = models.LCDM(lmax=5)
    base = models.parameters_Planck18(M)
    S0 = Setup(M, base)
    S1 = Setup(M, dict(base, h=0.70, Omega_c0=0.125 / 0.70**2, Omega_b0=0.0224 / 0.70**2))
    gm = GpuModel(M)
    for s, S in enumerate((S0, S1)):
        gm.upload_background(s, S.t, S.u, S.du, S.P)
    ks = 10 ** np.linspace(-1, 3.5, 12)
    calls = [   # (sets, k, saveat, kernel): the second call is smaller than the first
        (np.repeat([0, 1], len(ks)), np.tile(ks, 2), np.linspace(0.05, 2.5, 6), ALGO_WAVE),
        (np.array([1, 1, 0, 1, 0]), ks[[11, 8, 6, 3, 0]], np.zeros(0), ALGO_WAVE),
        (np.repeat([0, 1], len(ks)), np.tile(ks, 2), np.linspace(0.05, 2.5, 6), ALGO_WAVE2),
        (np.array([0, 1]), ks[[10, 2]], np.array([1.0]), ALGO_WAVE2),
    ]
    for c, (iset, kk, saveat, algo) in enumerate(calls):
        t0 = np.array([np.clip(min(1e-2 / k, 1e-4), (S0, S1)[s].t[0], (S0, S1)[s].t[-1]) for s, k in zip(iset, kk)])
        print(f"call {c}: {len(kk)} instances, {len(saveat)} saves, algo {algo}", flush=True)
        ref = batch(gm, "sbz_solve_batch", iset, kk, t0, saveat, algo)
        got = batch(gm, "sbz_solve_batch_noalloc", iset, kk, t0, saveat, algo)
        assert ref[0] == 0 and got[0] == 0, _lib.lib().sbz_last_error()
        assert np.all(ref[3] == 0)
        np.testing.assert_array_equal(got[1], ref[1])
        if len(saveat):
            np.testing.assert_array_equal(got[2], ref[2])
        np.testing.assert_array_equal(got[3], ref[3])
        np.testing.assert_array_equal(got[4], ref[4])
    # beyond the scratch array: an argument error, nothing launched
    n = 20000
    rc = batch(gm, "sbz_solve_batch_noalloc", np.zeros(n, np.int32), np.full(n, 1.0), np.full(n, 1e-4),
               np.linspace(0.1, 2.0, 30), ALGO_WAVE)[0]
    assert rc == 1
    L = _lib.lib()
    L.sbz_last_error.restype = ctypes.c_char_p
    assert b"scratch" in L.sbz_last_error()
    gm.close()
