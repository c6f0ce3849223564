"""CPU oracle of the CMB angular power spectra (TEST INFRASTRUCTURE ONLY).

Imported by tests/ and bench.py's cpu_baseline leg, and there only as the checker; the product
(symboltz.jl_amd) never imports it.

Restates `spectrum_cmb(modes, prob, jl[, ls])` (src/observables/angular.jl:260-359) step by step:
  grids          lingrid / cosgrid (src/utils.jl:270-285); τ grid = background knots ≥ τcut mapped
                 to the cosgrid fraction xs (angular.jl:280-292)
  k sampling     ChebyshevInterpolator (fourier.jl:419-458): second-kind Chebyshev points of
                 order n in y = f(k), f = tanh(k / 2000) when lensing is requested (angular.jl:266-272),
                 barycentric weights ±1 halved at both ends
  sources        source_grid on the Chebyshev nodes (fourier.jl:267-291) with
                 Ss = [k S_T, k² S_E, S_ψ] (angular.jl:296-298), last τ row zeroed (:299)
  k interp.      source_kinterp (fourier.jl:232-242, 520-532), C++ `cmbo_kinterp`
  LOS            los_integrate (angular.jl:109-185), C++ `cmbo_los`; Limber for ψ at l ≥ l_limber
                 (:309-311); Θ_T /= k, Θ_E *= √((l+2)(l+1)l(l-1)) / k² (:302-307)
  k integral     C_l = (2/π) ∫ dk k² P0 Θ_A Θ_B as the integral of a natural cubic spline through
                 (0, 0) and the fine samples (angular.jl:198-223; DataInterpolations CubicSpline),
                 here scipy's CubicSpline(bc_type="natural").integrate
  l interp.      natural cubic spline from the cached l to every requested l (angular.jl:347-359)
"""
import ctypes
import math

import numpy as np
from scipy.interpolate import CubicSpline

from . import oracle as _o

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_L = None
INT_MAX = 2**31 - 1


def lib():
    global _L
    if _L is None:
        if not _o.CMB_LIB.exists():
            _o.build()
        L = ctypes.CDLL(str(_o.CMB_LIB))
        L.cmbo_bessel_table.argtypes = [_ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.cmbo_kinterp.argtypes = [_dp, ctypes.c_int, ctypes.c_longlong, _dp, _dp, _dp, ctypes.c_int, _dp,
                                   ctypes.c_int]
        L.cmbo_los.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_int, _dp, _dp,
                               ctypes.c_int, ctypes.c_double, _ip, _dp, ctypes.c_int]
        _L = L
    return _L


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _threads():
    import os
    return min(16, os.cpu_count() or 1)


def lingrid(a, b, step=None, length=None):
    """utils.jl:270-278: `length` points from a to b inclusive, or ceil((b-a)/step)+1 of them."""
    if step is not None:
        length = int(math.ceil((b - a) / step)) + 1
    return np.linspace(a, b, length)


def cosgrid(a, b, length):
    # cospi(u) = sin(π (1/2 - u)) on [0, 1/2]: exact 1 and 0 at the ends, as Julia's cospi
    return a + (b - a) * (1 - np.sin(np.pi * (0.5 - lingrid(0.0, 0.5, length=length))))


def chebpoints(order, lb, ub):
    """Second-kind Chebyshev points cos(πj/n), j = 0..n, mapped to [lb, ub] (high to low)."""
    x = np.cos(np.pi * np.arange(order + 1) / order)
    return lb + (x + 1) * ((ub - lb) / 2)


class Cheb:
    def __init__(self, xmin, xmax, order, f=lambda k: k, finv=lambda y: y):
        self.f = f
        self.ys = chebpoints(order, f(xmin), f(xmax))
        self.xs = finv(self.ys)
        self.xs[-1], self.xs[0] = xmin, xmax
        w = np.where(np.arange(order + 1) % 2 == 0, 1.0, -1.0)
        w[0] /= 2
        w[-1] /= 2
        self.ws = w


def bessel_table(ls, xmax=None, dx=2 * math.pi / 15):
    """SphericalBesselCache grid and values: (dx_eff, y[n_x][n_l], dy[n_x][n_l]); n = trunc(xmax/dx)
    points from 0 to xmax inclusive plus a duplicate of the last (angular.jl:23-27)."""
    ls = np.ascontiguousarray(ls, dtype=np.int32)
    if xmax is None:
        xmax = 20.0 * float(ls[-1])
    n = int(xmax / dx)
    dxe = xmax / (n - 1)
    nx = n + 1
    y = np.zeros((nx, len(ls)))
    dy = np.zeros((nx, len(ls)))
    lib().cmbo_bessel_table(_p(ls, ctypes.c_int), len(ls), nx, dxe, _p(y), _p(dy), _threads())
    return dxe, y, dy


def kinterp(S_coarse, cheb, ks_fine):
    """S_coarse [n_c][...] → [n_f][...] by barycentric interpolation in y = f(k)."""
    S = np.ascontiguousarray(S_coarse, dtype=np.float64)
    m = int(np.prod(S.shape[1:]))
    yf = np.ascontiguousarray(cheb.f(np.asarray(ks_fine, dtype=np.float64)))
    out = np.zeros((len(yf),) + S.shape[1:])
    lib().cmbo_kinterp(_p(S), S.shape[0], m, _p(np.ascontiguousarray(cheb.ys)), _p(np.ascontiguousarray(cheb.ws)),
                       _p(yf), len(yf), _p(out), _threads())
    return out


def los(S, taus, ks, ls, table, l_limber):
    """S [n_k][n_τ][n_s] → Θ [n_s][n_k][n_l]."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    ls = np.ascontiguousarray(ls, dtype=np.int32)
    dxe, y, dy = table
    nk, nt, ns = S.shape
    lim = np.ascontiguousarray(l_limber, dtype=np.int32)
    out = np.zeros((ns, nk, len(ls)))
    lib().cmbo_los(_p(S), ns, _p(taus), nt, _p(ks), nk, _p(ls, ctypes.c_int), len(ls), _p(y), _p(dy), y.shape[0], dxe,
                   _p(lim, ctypes.c_int), _p(out), _threads())
    return out


def k_integral(ThA, ThB, P0, ks):
    """(2/π) ∫ dk k² P0 Θ_A Θ_B over the natural spline through (0, 0) and the samples; ThA [n_k][n_l]."""
    ks0 = np.concatenate([[0.0], ks])
    y = np.vstack([np.zeros((1, ThA.shape[1])), (2 / np.pi * ks**2 * P0)[:, None] * ThA * ThB])
    return CubicSpline(ks0, y, bc_type="natural").integrate(0.0, ks0[-1])


def spectrum_cmb(O, P0_fn, modes, ls, ls_fine=None, normalization="Dl", lensing=True, order=None, xs_len=300,
                 taucut=1e-2, l_limber=10, dk_tau0=math.pi, rtol=1e-5, atol=1e-5):
    """The oracle pipeline on an `oracle.Oracle` (its background knots are the τ source of the grid).

    Returns (spectra [n_l or n_l_fine][n_modes], details)."""
    if lensing:
        cheb = Cheb(1e-2, 1e4, 130 if order is None else order, f=lambda k: np.tanh(k / 2000.0),
                    finv=lambda y: 2000.0 * np.arctanh(y))
    else:
        cheb = Cheb(1e-2, 2e3, 60 if order is None else order)
    t = O.t
    tau0 = t[-1]
    ks_fine = lingrid(cheb.xs[-1], cheb.xs[0], step=dk_tau0 / tau0)
    tt = t[t >= taucut]
    taus = tt[0] + (tt[-1] - tt[0]) * cosgrid(0.0, 1.0, xs_len)
    kc = cheb.xs
    r = O.solve(kc, np.full(len(kc), t[0]), rtol=rtol, atol=atol, saveat=taus, sources=True)
    if np.any(r["retcode"] != 0):
        raise RuntimeError("oracle: a source mode failed")
    Sc = r["src"].copy()                      # [n_c][n_τ][3]
    Sc[:, :, 0] *= kc[:, None]
    Sc[:, :, 1] *= kc[:, None] ** 2
    Sc[:, -1, :] = 0.0
    Sf = kinterp(Sc, cheb, ks_fine)           # [n_f][n_τ][3]
    table = bessel_table(ls)
    Th = los(Sf, taus, ks_fine, ls, table, [INT_MAX, INT_MAX, l_limber])
    L = np.asarray(ls, dtype=np.float64)
    Th[0] /= ks_fine[:, None]
    Th[1] *= np.sqrt((L + 2) * (L + 1) * L * (L - 1))[None, :] / ks_fine[:, None] ** 2
    P0 = P0_fn(ks_fine)
    idx = {"T": 0, "E": 1, "ψ": 2}
    out = np.zeros((len(ls), len(modes)))
    for i, m in enumerate(modes):
        Cl = k_integral(Th[idx[m[0]]], Th[idx[m[-1]]], P0, ks_fine)
        out[:, i] = Cl * L * (L + 1) / (2 * np.pi) if normalization == "Dl" else Cl
    details = dict(taus=taus, ks_fine=ks_fine, ks_coarse=kc, Th=Th, S_coarse=Sc, stats=r["stats"])
    if ls_fine is not None:
        lf = np.asarray(ls_fine, dtype=np.float64)
        out = np.stack([CubicSpline(L, out[:, i], bc_type="natural")(lf) for i in range(len(modes))], axis=1)
    return out, details


def refine_queue(solve1, ks, f, finv, rtol, atol, its, nmaxks=1024):
    """source_grid_adaptive's refinement as the reference runs it without threads
    (fourier.jl:343-395): solve the initial points, queue their intervals, then take intervals
    first in first out, solve each midpoint, compare it with the linear interpolation from the
    ends (isapprox over the source vector at every τ of `its`) and queue (k2, k3) then (k1, k2).
    `solve1(k)` returns S [n_τ][n_S] of one wavenumber.  Returns (sorted k, S [n][n_τ][n_S])."""
    from collections import deque
    K = [float(k) for k in ks]
    S = [solve1(k) for k in K]
    q = deque((i - 1, i) for i in range(1, len(K)))
    while q:
        i1, i3 = q.popleft()
        if len(K) + 1 > nmaxks:
            raise RuntimeError("more than nmaxks points")
        k2 = finv((f(K[i1]) + f(K[i3])) / 2)
        S2 = solve1(k2)
        i2 = len(K)
        K.append(float(k2))
        S.append(S2)
        Si = (S[i1] + S[i3]) / 2
        bad = False
        for it in range(*its.indices(S2.shape[0])):
            d = np.linalg.norm(S2[it] - Si[it])
            if not d <= max(atol, rtol * max(np.linalg.norm(S2[it]), np.linalg.norm(Si[it]))):
                bad = True
        if bad:
            q.append((i2, i3))
            q.append((i1, i2))
    o = np.argsort(K, kind="stable")
    return np.array(K)[o], np.stack(S)[o]
