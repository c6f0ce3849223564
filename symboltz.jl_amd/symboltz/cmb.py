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
from .problem import solvebg, solvebg_sets, spectrum_primordial
from .solver import ALGO_AUTO, SAVE_SOURCES

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
    L.sbz_cmb_compute.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp, dp,
                                  dp, ctypes.c_int, ip, ip, ip, ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_float)]
    new_ctx = jl._ctx is None
    h = jl.context()
    S = np.ascontiguousarray(S_coarse, dtype=np.float64)
    n_c, n_t, n_s = S.shape
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    ks = np.ascontiguousarray(ks_fine, dtype=np.float64)
    yf = np.ascontiguousarray(kinterp.f(ks), dtype=np.float64)
    yc = np.ascontiguousarray(kinterp.ys, dtype=np.float64)
    wc = np.ascontiguousarray(kinterp.ws, dtype=np.float64)
    wk = np.ascontiguousarray(wk, dtype=np.float64)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    lim = np.ascontiguousarray(l_limber, dtype=np.int32)
    md = np.ascontiguousarray(modes, dtype=np.int32).ravel()
    n_m = len(md) // 2
    cl = np.zeros((n_m, len(jl.l)))
    th = np.zeros((n_s, len(ks), len(jl.l))) if want_theta else None
    ms = (ctypes.c_float * 5)()
    _lib.check(L.sbz_cmb_compute(h, _lib.ptr(S), n_c, n_s, _lib.ptr(taus), n_t, _lib.ptr(yc), _lib.ptr(wc),
                                 _lib.ptr(ks), _lib.ptr(yf), _lib.ptr(wk), len(ks), _lib.ptr(kinds, ctypes.c_int),
                                 _lib.ptr(lim, ctypes.c_int), _lib.ptr(md, ctypes.c_int), n_m, _lib.ptr(cl),
                                 _lib.ptr(th) if th is not None else None, ms), "sbz_cmb_compute")
    t = dict(zip(["bessel", "kinterp", "los", "limber", "kintegral"], list(ms)))
    t["bessel"] = jl.context_ms if new_ctx else 0.0     # the table is built once per cache
    return cl, th, t


def source_grid(prob, taus, ks, bg=None, ptopts=None, timing=None):
    """CMB sources (S_T, S_E, S_ψ) at every τ for every k, [n_k][n_τ][3] (fourier.jl:267-291), from
    one GPU launch; modes start at the background start (ptivini = -Inf)."""
    if bg is None:
        bg = solvebg(prob)
    return source_grid_sets([prob], [taus], ks, [bg], ptopts, timing)[0]


def source_grid_sets(probs, taus_sets, ks, bgs, ptopts=None, timing=None):
    """`source_grid` for several cosmologies of one model in ONE persistent launch: set s is solved
    at ks on its own time grid taus_sets[s]; returns [n_sets][n_k][n_τ][3]."""
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    key = probs[0].M.key()
    if any(p.M.key() != key for p in probs):
        raise ValueError("all problems of a batch must share one model")
    taus_sets = [np.asarray(t, dtype=np.float64) for t in taus_sets]
    n_t = len(taus_sets[0])
    for t, bg in zip(taus_sets, bgs):
        if len(t) != n_t:
            raise ValueError("every set needs the same number of τ points")
        if t.min() < bg.t[0] or t.max() > bg.t[-1]:
            raise ValueError("input τs and computed background solution have different timespans")
    gm = probs[0].gpu
    for s, bg in enumerate(bgs):
        gm.upload_background(s, bg.t, bg.u, bg.du, bg.P)
    # one saveat row per set held by the model (rows of sets outside this batch are never read)
    rows = np.repeat(taus_sets[0][None, :], gm.n_sets, axis=0)
    for s, t in enumerate(taus_sets):
        rows[s] = t
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    n_s, n_k = len(probs), len(ks)
    inst_set = np.repeat(np.arange(n_s, dtype=np.int32), n_k)
    inst_k = np.tile(ks, n_s)
    inst_t0 = np.repeat(np.array([bg.t[0] for bg in bgs]), n_k)
    plan = gm.plan(inst_set, inst_k, inst_t0, rtol=ptopts["reltol"], atol=ptopts["abstol"],
                   maxiters=ptopts["maxiters"], saveat=rows, save=SAVE_SOURCES,
                   algo=ptopts.get("algo", ALGO_AUTO))
    ms = plan.run()
    res = plan.fetch()
    src = plan.sources()
    plan.close()
    if timing is not None:
        timing["sources_ms"] = ms
        timing["steps"] = res["stats"][:, 0] + res["stats"][:, 1]    # accepted + rejected attempts
    bad = np.flatnonzero(res["retcode"] != 0)
    if len(bad):
        raise RuntimeError(f"Perturbation (mode k = {inst_k[bad[0]]}) solution failed with return code "
                           f"{_lib.RETCODE_NAMES[int(res['retcode'][bad[0]])]}")
    return src.reshape(n_s, n_k, n_t, 3)


_SRC_COL = {"ST": 0, "SE": 1, "Sψ": 2}


def refine_levels(solve, ks, f, finv, rtol, atol, its, nmaxks=1024, refine=True):
    """Bisection refinement of a k grid, generation by generation.  `solve(k[])` returns the source
    arrays [n][n_τ][n_S]; an interval (k1, k3) is split at k2 = f⁻¹((f(k1) + f(k3))/2) and both
    halves are queued again when, at some τ in `its`, |S(k2) - (S(k1) + S(k3))/2| exceeds
    max(atol, rtol max(|S(k2)|, |(S(k1) + S(k3))/2|)) (vector norms over the sources).
    Returns (k in insertion order, S [n][n_τ][n_S], number of solve calls)."""
    K = list(np.asarray(ks, dtype=np.float64))
    S = list(solve(np.asarray(K)))
    pending = [(i - 1, i) for i in range(1, len(K))] if refine else []
    calls = 1
    while pending:
        if len(K) + len(pending) > nmaxks:
            raise RuntimeError(f"Source function refinement needs more than {nmaxks} k-refinements. "
                               "Reduce refinement criteria.")
        mids = np.array([finv((f(K[i1]) + f(K[i3])) / 2) for i1, i3 in pending])
        Sm = solve(mids)
        calls += 1
        nxt = []
        for (i1, i3), k2, S2 in zip(pending, mids, Sm):
            i2 = len(K)
            K.append(float(k2))
            S.append(S2)
            Si = (S[i1] + S[i3]) / 2
            d = np.linalg.norm((S2 - Si)[its], axis=1)
            tol = np.maximum(atol, rtol * np.maximum(np.linalg.norm(S2[its], axis=1), np.linalg.norm(Si[its], axis=1)))
            if np.any(~(d <= tol)):
                nxt += [(i2, i3), (i1, i2)]
        pending = nxt
    return np.array(K), np.stack(S), calls


def source_grid_adaptive(prob, Ss, taus, ks, bg=None, ptopts=None, refine=True, ktransform=None, sort=True,
                         rtol=None, atol=0.0, nmaxks=1024, timing=None):
    """Sources on fixed τs and an adaptively refined k grid (fourier.jl:312-411).

    Every interval (k1, k3) of the grid is bisected in f(k) (ktransform = (f, f⁻¹), identity by
    default) until the source vector at the midpoint k2 matches the mean of its end values,
    `isapprox(S(k2), (S(k1) + S(k3))/2; rtol, atol)` with norm taken over the requested sources
    `Ss` (names among "ST", "SE", "Sψ") at each τ, the last τ excluded when there are several
    (1/χ terms diverge there).  rtol defaults to √eps when atol is 0 (Julia's isapprox).

    The reference refines interval by interval from a task queue; here every pending interval of
    a generation is bisected at once and all midpoints are solved in ONE GPU launch, which yields
    the same set of wavenumbers (each decision depends only on its interval).  More than `nmaxks`
    points is an error, as in the reference.  taus = None evaluates at τ0 only.
    Returns (ks [n], S [n_τ][n][len(Ss)]), sorted by k unless sort=False (then in insertion
    order: initial points, then each generation's midpoints in interval order)."""
    cols = [_SRC_COL[x] for x in Ss]
    ks = np.asarray(ks, dtype=np.float64)
    if len(ks) < 2:
        raise ValueError("Initial k-grid must have at least 2 values")
    f, finv = ktransform if ktransform is not None else ((lambda x: x), (lambda y: y))
    for kk in (ks.min(), ks.max()):
        if not np.isclose(kk, finv(f(kk))):
            raise ValueError("ktransform is not a tuple of inverse functions")
    if rtol is None:
        rtol = math.sqrt(np.finfo(np.float64).eps) if atol == 0 else 0.0
    if bg is None:
        bg = solvebg(prob)
    tsave = np.array([bg.tau0]) if taus is None else np.asarray(taus, dtype=np.float64)
    nt = len(tsave)
    its = slice(0, max(nt - 1, 1))

    def solve(kv):
        return source_grid(prob, tsave, kv, bg, ptopts, None)[:, :, cols]   # [n][n_τ][n_S]

    K, S, launches = refine_levels(solve, ks, f, finv, rtol, atol, its, nmaxks, refine)
    S = S.transpose(1, 0, 2)                   # [n_τ][n][n_S]
    if timing is not None:
        timing["launches"] = launches
    if sort:
        o = np.argsort(K, kind="stable")
        K, S = K[o], S[:, o]
    return K, S


_TEMPERATURE_UNITS = {"K": 1.0, "mK": 1e3, "uK": 1e6, "μK": 1e6}


def spectrum_cmb(modes, prob, jl, ls=None, normalization="Cl", kinterp=None, dk_tau0=math.pi, xs=None, taucut=1e-2,
                 l_limber=10, bgopts=None, ptopts=None, timing=None, details=None, bg=None, unit=None):
    """Angular power spectra C_l^{AB} for `modes` such as ["TT", "EE", "TE", "ψψ"] at jl.l (or, with
    `ls`, splined to every l in `ls`, whose extrema must equal those of jl.l), as
    `spectrum_cmb(modes, prob, jl[, ls]; normalization, unit)` (angular.jl:260-359).  Spectra are
    of dimensionless temperature fluctuations, or in unit² for unit "K", "mK" or "uK" (the factor
    (T0 in that unit)², angular.jl:318-324); normalization "Dl" returns l(l+1) C_l / 2π.
    Returns [n_l][n_modes].  `bg`: a background solution to use instead of solving one."""
    if unit is not None and unit not in _TEMPERATURE_UNITS:
        raise ValueError(f"Requested unit {unit} is not a temperature unit")
    dets = [] if details is not None else None
    out = spectrum_cmb_sets(modes, [prob], jl, ls, normalization, kinterp, dk_tau0, xs, taucut, l_limber, bgopts,
                            ptopts, timing, dets, bgs=None if bg is None else [bg])[0]
    if details is not None:
        details.update(dets[0])
    if unit is not None:
        from .models import P_T0
        out = out * (prob.P[P_T0] * _TEMPERATURE_UNITS[unit]) ** 2
    return out


def spectrum_cmb_sets(modes, probs, jl, ls=None, normalization="Cl", kinterp=None, dk_tau0=math.pi, xs=None,
                      taucut=1e-2, l_limber=10, bgopts=None, ptopts=None, timing=None, details=None, bgs=None):
    """`spectrum_cmb` for several cosmologies of one model: the source modes of every set run in one
    persistent launch (each set on its own τ grid), then the device C_l stage runs per set.
    Returns [n_sets][n_l][n_modes]; `details`, when a list, receives one dict per set.

    The backgrounds come from `bgs` when given, else from one batched GPU solve (several sets) or
    the host solve (one set).  The τ grid starts at the first background knot at or after τcut, as
    in the reference, so two background solvers whose step sequences differ in rounding give
    spectra that differ at the level of that grid's trapezoid error (~1e-4)."""
    if not isinstance(jl, SphericalBesselCache):
        raise TypeError("jl must be a SphericalBesselCache")
    modes = [str(m) for m in modes]
    for m in modes:
        if len(m) != 2 or m[0] not in _KIND or m[1] not in _KIND:
            raise ValueError(f"Unknown CMB power spectrum mode {m}")
    if normalization not in ("Cl", "Dl"):
        raise ValueError(f"Normalization {normalization} is not Cl or Dl")
    if ls is not None:
        ls = np.asarray(ls)
        if (ls.min(), ls.max()) != (jl.l.min(), jl.l.max()):
            raise ValueError(f"jl.l and ls have different extrema {(jl.l.min(), jl.l.max())} != "
                             f"{(ls.min(), ls.max())}")
    joined = "".join(modes)
    present = [s for s in "TEψ" if s in joined]
    if kinterp is None:
        if "ψ" in present:
            kinterp = ChebyshevInterpolator(1e-2, 1e4, 130, f=fk_tanh, finv=fk_inv_tanh)
        else:
            kinterp = ChebyshevInterpolator(1e-2, 2e3, 60)
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    xs = cosgrid(0.0, 1.0, length=300) if xs is None else np.asarray(xs, dtype=np.float64)
    if xs[0] != 0 or xs[-1] != 1:
        raise ValueError(f"xs begins with {xs[0]} and ends with {xs[-1]}, but should span [0, 1]")
    if bgs is not None:
        if len(bgs) != len(probs):
            raise ValueError("one background solution per problem")
    elif len(probs) > 1:   # every set's background at once on the GPU
        bgs = solvebg_sets(probs, bgopts["reltol"], bgopts["abstol"])
    else:
        bgs = [solvebg(probs[0], bgopts["reltol"], bgopts["abstol"])]
    taus_sets = []
    for bg in bgs:
        t = bg.t[bg.t >= taucut]
        taus = t[0] + (t[-1] - t[0]) * xs
        taus_sets.append(taus)
    kc = kinterp.xs
    src = source_grid_sets(probs, taus_sets, kc, bgs, ptopts, timing)   # [n_sets][n_c][n_τ][3]
    cols = [_KIND[s] for s in present]
    kinds = [_KIND[s] for s in present]
    lim = [l_limber if s == "ψ" else INT_MAX for s in present]
    pairs = [(present.index(m[0]), present.index(m[1])) for m in modes]
    lc = np.asarray(jl.l, dtype=np.float64)
    outs = []
    for si, (prob, bg, taus) in enumerate(zip(probs, bgs, taus_sets)):
        ks_fine = lingrid(kinterp.minimum(), kinterp.maximum(), step=dk_tau0 / bg.tau0)
        if jl.x[-1] < ks_fine[-1] * taus[-1]:
            raise ValueError("the Bessel cache ends below k_max τ0")
        S = src[si][:, :, cols].copy()
        for i, s in enumerate(present):
            if s == "T":
                S[:, :, i] *= kc[:, None]
            elif s == "E":
                S[:, :, i] *= kc[:, None] ** 2
        S[:, -1, :] = 0.0                      # χ = 0: weighted by j_l(0) = 0, may be non-finite
        P0 = spectrum_primordial(ks_fine, prob)
        q = spline_integral_weights(np.concatenate([[0.0], ks_fine]))[1:]
        wk = q * (2 / np.pi) * ks_fine**2 * P0
        cl, th, ms = _cmb_spectra_gpu(S, taus, kinterp, ks_fine, wk, jl, kinds, lim, pairs,
                                      want_theta=details is not None)
        if timing is not None:
            for k, v in ms.items():
                timing[f"{k}_ms"] = timing.get(f"{k}_ms", 0.0) + v
            timing["n_k_fine"] = len(ks_fine)
        if details is not None:
            details.append(dict(taus=taus, ks_fine=ks_fine, ks_coarse=kc, Th=th, S_coarse=S, present=present))
        out = cl.T.copy()
        if normalization == "Dl":
            out *= (lc * (lc + 1) / (2 * np.pi))[:, None]
        if ls is not None:
            lf = ls.astype(np.float64)
            out = np.stack([CubicSpline(lc, out[:, i], bc_type="natural")(lf) for i in range(out.shape[1])], axis=1)
        outs.append(out)
    return np.stack(outs)
