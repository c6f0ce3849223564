"""Line-of-sight integration and the C_l k-integral.

`los_integrate` runs on the GPU (sbz_los_integrate in libsbz_hip.so) with the semantics of the
reference's `los_integrate` (src/observables/angular.jl:109-185, no Limber branch):
I_l(k) = Σ_τ w_τ S(τ, k) j_l(k(τ0 - τ)), trapezoid weights, τ0 = τs[-1].
`spectrum_cmb_from_los` is the k-integral C_l = (2/π) ∫ dk k² P0(k) Θ_l^A Θ_l^B over a cubic
spline with a (0, 0) point prepended (angular.jl:198-223).
"""
import ctypes

import numpy as np
from scipy.interpolate import CubicSpline

from . import _lib


def los_integrate(Ss, ls, taus, ks, jl, timing=None):
    """Ss[nτ, nk] → I[nk, nl] on the GPU."""
    L = _lib.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.sbz_los_integrate.argtypes = [dp, ctypes.c_longlong, ctypes.c_longlong, dp, ctypes.c_int, dp, ctypes.c_int,
                                    dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp,
                                    ctypes.POINTER(ctypes.c_float)]
    Ss = np.ascontiguousarray(Ss, dtype=np.float64)
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    if not np.array_equal(np.asarray(ls), jl.l):
        raise ValueError("ls must match the l-values stored in the Bessel cache")
    if Ss.shape != (len(taus), len(ks)):
        raise ValueError(f"Ss has shape {Ss.shape}, expected ({len(taus)}, {len(ks)})")
    if not np.all(np.isfinite(Ss)):
        raise ValueError("non-finite source values")
    if not (np.all(ks > 0) and np.all(np.isfinite(ks))):
        raise ValueError("wavenumbers must be positive and finite")
    nl = len(jl.l)
    out = np.zeros((len(ks), nl))
    ms = ctypes.c_float(0.0)
    _lib.check(L.sbz_los_integrate(_lib.ptr(Ss), 1, len(ks), _lib.ptr(taus), len(taus), _lib.ptr(ks), len(ks),
                                   _lib.ptr(jl.y), _lib.ptr(jl.dy), len(jl.x), jl.dx, nl, _lib.ptr(out),
                                   ctypes.byref(ms)), "sbz_los_integrate")
    if timing is not None:
        timing["kernel_ms"] = ms.value
    return out


def spectrum_cmb_from_los(ThA, ThB, P0s, ls, ks, normalization="Cl"):
    ks0 = np.concatenate([[0.0], ks])
    Cl = np.zeros(len(ls))
    for il in range(len(ls)):
        y = np.concatenate([[0.0], 2 / np.pi * ks**2 * P0s * ThA[:, il] * ThB[:, il]])
        Cl[il] = CubicSpline(ks0, y, bc_type="natural").integrate(ks0[0], ks0[-1])
    if normalization == "Dl":
        ls = np.asarray(ls, dtype=np.float64)
        return Cl * ls * (ls + 1) / (2 * np.pi)
    return Cl
