"""ctypes binding of libsbz_hip.so (the C ABI declared in include/sbz.h).

The library is built in-tree by `symboltz.build.build_library()` (driven by the repository's
`__graft_entry__.build()`); this module only loads it.  A missing library is an error: there is
no CPU fallback for any entry point.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent          # .../symboltz.jl_amd
LIB_PATH = PKG_ROOT / "lib" / "libsbz_hip.so"

SBZ_OK = 0
RETCODE_NAMES = {0: "Success", 1: "MaxIters", 2: "DtLessThanMin", 3: "Unstable", 4: "InitialFailure",
                 5: "NotRun"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

_lib = None


class SbzError(RuntimeError):
    pass


def ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def lib():
    """Load libsbz_hip.so (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("SBZ_LIB", str(LIB_PATH)))
    if not path.exists():
        raise SbzError(f"libsbz_hip.so not built at {path}; run __graft_entry__.build()")
    L = ctypes.CDLL(str(path))
    L.sbz_last_error.restype = ctypes.c_char_p
    L.sbz_version.restype = ctypes.c_int
    L.sbz_device_count.argtypes = [_ip]
    L.sbz_background_solve.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp,
                                       _ip, _ip]
    _declare_solver(L)
    _lib = L
    return L


def _declare_solver(L):
    vp = ctypes.c_void_p
    if hasattr(L, "sbz_model_create"):
        L.sbz_model_create.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(vp)]
        L.sbz_model_destroy.argtypes = [vp]
        L.sbz_bg_upload.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.sbz_solve_batch.argtypes = [vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, _dp, ctypes.c_int,
                                      _dp, _ip, ctypes.c_void_p]


def check(status, what=""):
    if status != SBZ_OK:
        msg = lib().sbz_last_error().decode(errors="replace")
        raise SbzError(f"{what}: status {status}: {msg}")


def device_count():
    n = ctypes.c_int(0)
    check(lib().sbz_device_count(ctypes.byref(n)), "sbz_device_count")
    return n.value


def background_solve(P, rtol=1e-7, atol=1e-7, max_knots=20000):
    """Run the host background solve; returns (P_out, t, u, du, n_rejected)."""
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    t = np.zeros(max_knots)
    u = np.zeros((max_knots, 5))
    du = np.zeros((max_knots, 5))
    nk = ctypes.c_int(0)
    nrej = ctypes.c_int(0)
    st = lib().sbz_background_solve(ptr(P), rtol, atol, max_knots, ptr(t), ptr(u), ptr(du), ctypes.byref(nk),
                                    ctypes.byref(nrej))
    check(st, "sbz_background_solve")
    n = nk.value
    return P, t[:n].copy(), u[:n].copy(), du[:n].copy(), nrej.value


BG_STATUS = {0: "Success", 1: "MaxIters", 2: "DtLessThanMin", 3: "MaxKnots", 4: "NoEnd"}


def background_solve_batch(Ps, rtol=1e-7, atol=1e-7, max_knots=2048):
    """The background solve of every row of Ps [n][SBZ_NPAR] on the GPU, one thread per set.
    Returns (P_out [n][NPAR], list of (t, u, du) per set, n_rejected [n], status [n], kernel_ms)."""
    L = lib()
    ip = ctypes.POINTER(ctypes.c_int)
    L.sbz_background_solve_batch.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             _dp, _dp, _dp, ip, ip, ip, ctypes.POINTER(ctypes.c_float)]
    P = np.ascontiguousarray(Ps, dtype=np.float64).copy()
    n = P.shape[0]
    t = np.empty((n, max_knots))          # only each row's used prefix is written
    u = np.empty((n, max_knots, 5))
    du = np.empty((n, max_knots, 5))
    nk = np.zeros(n, dtype=np.int32)
    nr = np.zeros(n, dtype=np.int32)
    stt = np.zeros(n, dtype=np.int32)
    ms = ctypes.c_float(0.0)
    check(L.sbz_background_solve_batch(ptr(P), n, rtol, atol, max_knots, ptr(t), ptr(u), ptr(du),
                                       ptr(nk, ctypes.c_int), ptr(nr, ctypes.c_int), ptr(stt, ctypes.c_int),
                                       ctypes.byref(ms)), "sbz_background_solve_batch")
    knots = [(t[i, :nk[i]].copy(), u[i, :nk[i]].copy(), du[i, :nk[i]].copy()) for i in range(n)]
    return P, knots, nr, stt, ms.value
