"""CMB angular power spectra: `spectrum_cmb` (src/observables/angular.jl:198-359) on the GPU.

The pipeline of the reference, stage by stage:
  1. background solve; τ grid = the background knots ≥ τcut mapped to the fraction grid xs
     (default cosgrid(0, 1; length = 300)) (angular.jl:276-292)
  2. the perturbation modes at the nodes of a ChebyshevInterpolator in k (fourier.jl:419-458;
     tanh-stretched 130th order on [1e-2, 1e4] when lensing is requested, linear 60th order on
     [1e-2, 2e3] otherwise, angular.jl:266-272), solved on the GPU with the CMB sources saved at
     every τ (source_grid, fourier.jl:267-291), Ss = [k S_T, k² S_E, S_ψ] (angular.jl:294-299)
  3. on the device (`sbz_cmb_spectra`): barycentric interpolation to the fine grid
     lingrid(kmin, kmax; step = Δkτ0/τ0), the Bessel table, line-of-sight integrals with the Limber
     approximation for S_ψ at l ≥ l_limber, the T and E factors, and the k integral
  4. the natural cubic spline from the cached l to every requested l (angular.jl:347-359).

The k integral of the reference is the exact integral of a natural cubic spline through (0, 0)
and the fine samples (DataInterpolations.CubicSpline).  That integral is linear in the samples,
so it is applied here as fixed quadrature weights q (`spline_integral_weights`) that depend on
the k grid only: C_l = Σ_k q_k (2/π) k² P0(k) Θ_A Θ_B.
"""
import ctypes
import math

import numpy as np
from scipy.interpolate import CubicSpline

from . import _lib
from .bessel import SphericalBesselCache
from .problem import solvebg, spectrum_primordial
from .solver import SAVE_SOURCES

INT_MAX = 2**31 - 1
_KIND = {"T": 0, "E": 1, "ψ": 2}


def lingrid(a, b, step=None, length=None):
    """`length` points from a to b inclusive, or ceil((b - a)/step) + 1 of them (utils.jl:270-278)."""
    if a >= b:
        raise ValueError("Right grid interval must be higher than left")
    if (step is None) == (length is None):
        raise ValueError("Must provide exactly one of step and length")
    if step is not None:
        length = int(math.ceil((b - a) / step)) + 1
    return np.linspace(a, b, length)


def cosgrid(a, b, step=None, length=None):
    """a + (b - a)(1 - cos(π u)) on u = lingrid(0, 1/2) (utils.jl:285)."""
    u = lingrid(0.0, 0.5, step=None if step is None else step / math.pi, length=length)
    return a + (b - a) * (1 - np.sin(np.pi * (0.5 - u)))    # cospi(u), exact at u = 0 and 1/2


def fk_tanh(k, k0=2000.0):
    return np.tanh(np.asarray(k) / k0)


def fk_inv_tanh(y, k0=2000.0):
    return k0 * np.arctanh(np.asarray(y))


class ChebyshevInterpolator:
    """Barycentric interpolation on second-kind Chebyshev points in y = f(x) (fourier.jl:419-458,
    520-532).  `xs`/`ys` run from high to low x; weights ±1, halved at both ends."""

    def __init__(self, xmin, xmax, order, f=None, finv=None):
        if not xmax > xmin:
            raise ValueError(f"Interval {(xmin, xmax)} is not sorted")
        self.f = f if f is not None else (lambda x: np.asarray(x, dtype=np.float64))
        if f is None:
            finv = self.f
        elif finv is None:
            raise ValueError("f needs its inverse finv")
        ymin, ymax = float(self.f(xmin)), float(self.f(xmax))
        j = np.arange(order + 1)
        self.ys = ymin + (np.cos(np.pi * j / order) + 1) * ((ymax - ymin) / 2)
        if not np.all(np.diff(self.ys) < 0):
            raise ValueError("Domain transformation f(x) is not monotonically increasing")
        xs = np.asarray(finv(self.ys), dtype=np.float64)
        if not (np.isclose(xs[-1], xmin) and np.isclose(xs[0], xmax)):
            raise ValueError("f(x) and f⁻¹(x) are not inverses")
        xs[-1], xs[0] = xmin, xmax
        self.xs = xs
        self.ws = np.where(j % 2 == 0, 1.0, -1.0)
        self.ws[0] /= 2
        self.ws[-1] /= 2

    def minimum(self):
        return self.xs[-1]

    def maximum(self):
        return self.xs[0]

    def __len__(self):
        return len(self.xs)

    def __call__(self, fvals, x):
        """Host evaluation at x of the interpolant through (xs, fvals)."""
        y = float(self.f(x))
        d = y - self.ys
        hit = np.flatnonzero(d == 0)
        if len(hit):
            return fvals[hit[0]]
        t = self.ws / d
        return np.dot(t, fvals) / np.sum(t)


def spline_integral_weights(x):
    """Weights q with Σ q_i y_i = ∫ s(x) dx over [x_0, x_N] for the natural cubic spline s through
    (x_i, y_i).  With h_i = x_{i+1} - x_i and second derivatives M (M_0 = M_N = 0),
      ∫ s = Σ_i h_i (y_i + y_{i+1})/2 - Σ_i h_i³ (M_i + M_{i+1})/24,
    and T M_int = B y with T_ii = 2(h_{i-1} + h_i), T_{i,i±1} = h_i / h_{i-1} (symmetric) and
    B y = 6 [(y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1}].  Hence q = a - Bᵀ T⁻¹ c, with
    a_i = (h_{i-1} + h_i)/2 and c_i = (h_{i-1}³ + h_i³)/24 for the interior points."""
    from scipy.linalg import solve_banded
    x = np.asarray(x, dtype=np.float64)
    N = len(x) - 1
    h = np.diff(x)
    a = np.zeros(N + 1)
    a[:-1] += h / 2
    a[1:] += h / 2
    if N < 2:
        return a
    hm, hp = h[:-1], h[1:]                        # h_{i-1}, h_i for interior i = 1..N-1
    c = (hm**3 + hp**3) / 24
    ab = np.zeros((3, N - 1))
    ab[0, 1:] = hp[:-1]
    ab[1] = 2 * (hm + hp)
    ab[2, :-1] = hp[:-1]
    z = solve_banded((1, 1), ab, c)
    q = a.copy()
    # Bᵀ z: row i of B has 6/h_{i-1} at i-1, -6/h_{i-1} - 6/h_i at i, 6/h_i at i+1
    q[:-2] -= 6 / hm * z
    q[1:-1] -= (-6 / hm - 6 / hp) * z
    q[2:] -= 6 / hp * z
    return q


def _cmb_spectra_gpu(S_coarse, taus, kinterp, ks_fine, wk, jl, kinds, l_limber, modes, want_theta=False):
    L = _lib.lib()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.sbz_cmb_spectra.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int,
                                  ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, ip, ip, ip, ctypes.c_int, dp, dp,
                                  ctypes.POINTER(ctypes.c_float)]
    S = np.ascontiguousarray(S_coarse, dtype=np.float64)
    n_c, n_t, n_s = S.shape
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    ks = np.ascontiguousarray(ks_fine, dtype=np.float64)
    yf = np.ascontiguousarray(kinterp.f(ks), dtype=np.float64)
    yc = np.ascontiguousarray(kinterp.ys, dtype=np.float64)
    wc = np.ascontiguousarray(kinterp.ws, dtype=np.float64)
    wk = np.ascontiguousarray(wk, dtype=np.float64)
    ls = np.ascontiguousarray(jl.l, dtype=np.int32)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    lim = np.ascontiguousarray(l_limber, dtype=np.int32)
    md = np.ascontiguousarray(modes, dtype=np.int32).ravel()
    n_m = len(md) // 2
    cl = np.zeros((n_m, len(ls)))
    th = np.zeros((n_s, len(ks), len(ls))) if want_theta else None
    ms = (ctypes.c_float * 5)()
    _lib.check(L.sbz_cmb_spectra(_lib.ptr(S), n_c, n_s, _lib.ptr(taus), n_t, _lib.ptr(yc), _lib.ptr(wc), _lib.ptr(ks),
                                 _lib.ptr(yf), _lib.ptr(wk), len(ks), _lib.ptr(ls, ctypes.c_int), len(ls), jl.n_x,
                                 jl.dx, _lib.ptr(kinds, ctypes.c_int), _lib.ptr(lim, ctypes.c_int),
                                 _lib.ptr(md, ctypes.c_int), n_m, _lib.ptr(cl),
                                 _lib.ptr(th) if th is not None else None, ms), "sbz_cmb_spectra")
    return cl, th, dict(zip(["bessel", "kinterp", "los", "limber", "kintegral"], list(ms)))


def source_grid(prob, taus, ks, bg=None, ptopts=None, timing=None):
    """CMB sources (S_T, S_E, S_ψ) at every τ for every k, [n_k][n_τ][3] (fourier.jl:267-291), from
    one GPU launch; modes start at the background start (ptivini = -Inf)."""
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    if bg is None:
        bg = solvebg(prob)
    taus = np.asarray(taus, dtype=np.float64)
    if taus.min() < bg.t[0] or taus.max() > bg.t[-1]:
        raise ValueError("input τs and computed background solution have different timespans")
    gm = prob.gpu
    gm.upload_background(0, bg.t, bg.u, bg.du, bg.P)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, np.full(len(ks), bg.t[0]), rtol=ptopts["reltol"],
                   atol=ptopts["abstol"], maxiters=ptopts["maxiters"], saveat=taus, save=SAVE_SOURCES)
    ms = plan.run()
    res = plan.fetch()
    src = plan.sources()
    plan.close()
    if timing is not None:
        timing["sources_ms"] = ms
        timing["steps"] = res["stats"][:, 0]
    bad = np.flatnonzero(res["retcode"] != 0)
    if len(bad):
        raise RuntimeError(f"Perturbation (mode k = {ks[bad[0]]}) solution failed with return code "
                           f"{_lib.RETCODE_NAMES[int(res['retcode'][bad[0]])]}")
    return src


def spectrum_cmb(modes, prob, jl, ls=None, normalization="Cl", kinterp=None, dk_tau0=math.pi, xs=None, taucut=1e-2,
                 l_limber=10, bgopts=None, ptopts=None, timing=None, details=None):
    """Angular power spectra C_l^{AB} for `modes` such as ["TT", "EE", "TE", "ψψ"] at jl.l (or, with
    `ls`, splined to every l in `ls`, whose extrema must equal those of jl.l), as
    `spectrum_cmb(modes, prob, jl[, ls]; normalization)` (angular.jl:260-359).  Spectra are of
    dimensionless temperature fluctuations; normalization "Dl" returns l(l+1) C_l / 2π.
    Returns [n_l][n_modes]."""
    if not isinstance(jl, SphericalBesselCache):
        raise TypeError("jl must be a SphericalBesselCache")
    modes = [str(m) for m in modes]
    for m in modes:
        if len(m) != 2 or m[0] not in _KIND or m[1] not in _KIND:
            raise ValueError(f"Unknown CMB power spectrum mode {m}")
    if normalization not in ("Cl", "Dl"):
        raise ValueError(f"Normalization {normalization} is not Cl or Dl")
    joined = "".join(modes)
    present = [s for s in "TEψ" if s in joined]
    if kinterp is None:
        if "ψ" in present:
            kinterp = ChebyshevInterpolator(1e-2, 1e4, 130, f=fk_tanh, finv=fk_inv_tanh)
        else:
            kinterp = ChebyshevInterpolator(1e-2, 2e3, 60)
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    bg = solvebg(prob, bgopts["reltol"], bgopts["abstol"])
    tau0 = bg.tau0
    ks_fine = lingrid(kinterp.minimum(), kinterp.maximum(), step=dk_tau0 / tau0)
    t = bg.t[bg.t >= taucut]
    xs = cosgrid(0.0, 1.0, length=300) if xs is None else np.asarray(xs, dtype=np.float64)
    if xs[0] != 0 or xs[-1] != 1:
        raise ValueError(f"xs begins with {xs[0]} and ends with {xs[-1]}, but should span [0, 1]")
    taus = t[0] + (t[-1] - t[0]) * xs
    if jl.x[-1] < ks_fine[-1] * taus[-1]:
        raise ValueError("the Bessel cache ends below k_max τ0")

    src = source_grid(prob, taus, kinterp.xs, bg, ptopts, timing)   # [n_c][n_τ][3]
    kc = kinterp.xs
    cols = [_KIND[s] for s in present]
    S = src[:, :, cols].copy()
    for i, s in enumerate(present):
        if s == "T":
            S[:, :, i] *= kc[:, None]
        elif s == "E":
            S[:, :, i] *= kc[:, None] ** 2
    S[:, -1, :] = 0.0                          # χ = 0: weighted by j_l(0) = 0, may be non-finite
    kinds = [_KIND[s] for s in present]
    lim = [l_limber if s == "ψ" else INT_MAX for s in present]
    P0 = spectrum_primordial(ks_fine, prob)
    q = spline_integral_weights(np.concatenate([[0.0], ks_fine]))[1:]
    wk = q * (2 / np.pi) * ks_fine**2 * P0
    pairs = [(present.index(m[0]), present.index(m[1])) for m in modes]
    cl, th, ms = _cmb_spectra_gpu(S, taus, kinterp, ks_fine, wk, jl, kinds, lim, pairs,
                                  want_theta=details is not None)
    if timing is not None:
        timing.update({f"{k}_ms": v for k, v in ms.items()})
        timing["n_k_fine"] = len(ks_fine)
    if details is not None:
        details.update(taus=taus, ks_fine=ks_fine, ks_coarse=kc, Th=th, S_coarse=S, present=present)
    lc = np.asarray(jl.l, dtype=np.float64)
    out = cl.T.copy()
    if normalization == "Dl":
        out *= (lc * (lc + 1) / (2 * np.pi))[:, None]
    if ls is not None:
        ls = np.asarray(ls)
        if (ls.min(), ls.max()) != (jl.l.min(), jl.l.max()):
            raise ValueError(f"jl.l and ls have different extrema {(jl.l.min(), jl.l.max())} != "
                             f"{(ls.min(), ls.max())}")
        lf = ls.astype(np.float64)
        out = np.stack([CubicSpline(lc, out[:, i], bc_type="natural")(lf) for i in range(out.shape[1])], axis=1)
    return out
