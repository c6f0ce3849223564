"""Host-side mirror of SymBoltz's problem/solve API for the perturbation stage.

Names, argument meaning and error behaviour follow the reference:
  CosmologyProblem(M, pars)            src/solve.jl:129-236 (compile step: here the emitted model)
  remake / parameter_updater            src/solve.jl:248-307
  solvebg(prob)                         src/solve.jl:427-435 (host C++ background solver)
  solve(prob, ks; ptopts, ptivini)      src/solve.jl:380-402 → solvept :543-569 on the GPU
  spectrum_matter(prob, ks)             src/observables/fourier.jl:79-101
  spectrum_primordial(k, prob)          src/observables/fourier.jl:14-37
  issuccess(sol)                        src/solve.jl:604-606
Per-mode failures do not raise: they warn "Perturbation (mode k = …) solution failed with return
code …" (solve.jl:557-564) and are visible in the retcodes.  Argument errors raise.

Wavenumbers are dimensionless (units of H0/c) unless `k_unit="1/Mpc"` is given, as the
reference converts Unitful wavenumbers with k_dimensionless (src/constants.jl:47-50).
"""
import math
import warnings

import numpy as np

from . import _lib
from . import constants as C
from .models import (P_AS, P_KAPPA0, P_KPIVOT, P_NS, P_TAU0, P_TAUREC, P_H, Model, make_params)

RETCODES = _lib.RETCODE_NAMES


class BackgroundSolution:
    """Background solve: knots t, unknowns u (a, X_H+, X_He+, ΔT, _κ), derivatives du, parameters P."""

    def __init__(self, P, t, u, du, n_rejected):
        self.P, self.t, self.u, self.du, self.n_rejected = P, t, u, du, n_rejected
        self.retcode = "Success"

    @property
    def tau0(self):
        return float(self.P[P_TAU0])

    @property
    def kappa0(self):
        return float(self.P[P_KAPPA0])

    @property
    def taurec(self):
        return float(self.P[P_TAUREC])


class CosmologyProblem:
    """A model plus parameter values.  The perturbation system is emitted and compiled for the GPU
    on first use (the reference compiles it in the constructor, src/solve.jl:215-227)."""

    def __init__(self, M, pars, ivspan=(1e-6, 100.0)):
        if not isinstance(M, Model):
            raise TypeError("M must be a symboltz Model")
        self.M = M
        self.pars = dict(pars)
        self.ivspan = ivspan
        self.P = make_params(M, self.pars, tau_bg0=ivspan[0])
        self._gpu = None

    @property
    def gpu(self):
        if self._gpu is None:
            self._gpu = _gpu_model(self.M)
        return self._gpu

    def __repr__(self):
        return f"Cosmology problem for model {self.M.name}: perturbations {self.M.n_unknowns} unknowns, 5 splines"


_gpu_models = {}


def _gpu_model(M):
    from .solver import GpuModel
    key = M.key()
    if key not in _gpu_models:
        _gpu_models[key] = GpuModel(M)
    return _gpu_models[key]


def remake(prob, pars):
    """New problem with some parameters replaced (src/solve.jl:248-264)."""
    p = dict(prob.pars)
    p.update(pars)
    q = CosmologyProblem(prob.M, p, prob.ivspan)
    q._gpu = prob._gpu
    return q


def parameter_updater(prob, names):
    """Return θ ↦ remake(prob, dict(zip(names, θ))) (src/solve.jl:272-307)."""
    names = list(names)

    def updater(values):
        if isinstance(values, dict):
            values = [values[n] for n in names]
        return remake(prob, dict(zip(names, values)))
    return updater


def solvebg(prob, reltol=1e-7, abstol=1e-7, max_knots=20000):
    P, t, u, du, nrej = _lib.background_solve(prob.P, reltol, abstol, max_knots)
    return BackgroundSolution(P, t, u, du, nrej)


GPU_BACKGROUND_MIN_SETS = 256


def solvebg_sets(probs, reltol=1e-7, abstol=1e-7, max_knots=2048, device=None, timing=None):
    """`solvebg` for many problems at once.  On the GPU (`device=True`) one thread integrates one
    parameter set: a thread is ~15x slower than a host core on this serial solve, so the GPU pays
    from a few hundred sets on; `device=None` picks the GPU for at least GPU_BACKGROUND_MIN_SETS sets
    and host threads (16) otherwise.  A set whose solve fails raises with the reference's wording
    ("Background solution failed")."""
    if device is None:
        device = len(probs) >= GPU_BACKGROUND_MIN_SETS
    if not device:
        import time
        from concurrent.futures import ThreadPoolExecutor   # the host solves release the GIL
        t = time.perf_counter()
        with ThreadPoolExecutor(max_workers=min(16, len(probs))) as ex:
            out = list(ex.map(lambda p: solvebg(p, reltol, abstol), probs))
        if timing is not None:
            timing["background_ms"] = (time.perf_counter() - t) * 1e3
        return out
    Ps = np.stack([p.P for p in probs])
    P, knots, nrej, status, ms = _lib.background_solve_batch(Ps, reltol, abstol, max_knots)
    if timing is not None:
        timing["background_ms"] = ms
    bad = np.flatnonzero(status != 0)
    if len(bad):
        raise RuntimeError(f"Background solution failed for parameter set {bad[0]} with return code "
                           f"{_lib.BG_STATUS.get(int(status[bad[0]]), status[bad[0]])}")
    return [BackgroundSolution(P[i], *knots[i], int(nrej[i])) for i in range(len(probs))]


BG_UNKNOWNS = {"a": 0, "XH": 1, "XHe": 2, "dT": 3, "kappa": 4}


def _hermite(t, y, dy, x):
    i = np.clip(np.searchsorted(t, x, side="right") - 1, 0, len(t) - 2)
    h = t[i + 1] - t[i]
    s = (x - t[i]) / h
    return ((2 * s**3 - 3 * s**2 + 1) * y[i] + (s**3 - 2 * s**2 + s) * h * dy[i] + (-2 * s**3 + 3 * s**2) * y[i + 1]
            + (s**3 - s**2) * h * dy[i + 1])


class CosmologySolution:
    """Background knots plus the perturbation modes of one `solve` call.

    Calling it mirrors `sol(is, τs[, ks])` (src/solve.jl:720-787):
    - background unknowns by name ("a", "XH", "XHe", "dT", "kappa") read from the cubic Hermite
      spline of the knots;
    - perturbation unknowns by index into the state vector at (τ, k), linear in log k between the
      two neighbouring solved modes.

    The modes' dense output at arbitrary τ comes from the GPU: the needed modes are integrated again
    with `saveat = τs`. The integration is deterministic, so they take the same steps as in the
    original solve and return its cubic Hermite dense output."""

    def __init__(self, prob, bg, ks, pts, t0=None, ptopts=None):
        self.prob, self.bg, self.ks, self.pts = prob, bg, ks, pts
        self._t0, self._ptopts = t0, ptopts

    def _modes_at(self, modes, taus):
        """States [n_modes][n_τ][n] of solved modes at the times τs.  Times must lie in the solved
        timespan [τ_bg0, τ0] (the states before a mode's start time are its initial state);
        a mode that failed has NaN at the times it never reached, with the reference's warning."""
        t_lo, t_hi = self.bg.t[0], self.bg.t[-1]
        if np.any(~(taus >= t_lo)) or np.any(~(taus <= t_hi)):
            raise ValueError(f"Requested times must lie in the solved timespan [{t_lo}, {t_hi}]")
        gm = self.prob.gpu
        gm.upload_background(0, self.bg.t, self.bg.u, self.bg.du, self.bg.P)
        o = self._ptopts
        order = np.argsort(taus, kind="stable")             # saveat must ascend
        plan = gm.plan(np.zeros(len(modes), dtype=np.int32), self.ks[modes], self._t0[modes], rtol=o["reltol"],
                       atol=o["abstol"], maxiters=o["maxiters"], saveat=taus[order])
        try:
            plan.run()
            r = plan.fetch(save=True)
        finally:
            plan.close()
        for j, rc in enumerate(r["retcode"]):
            if rc != 0:
                warnings.warn(f"Perturbation (mode k = {self.ks[modes[j]]}) solution failed with return code "
                              f"{RETCODES[int(rc)]}; its states past the failure are NaN.")
        u = r["usave"]                                      # [n_modes][n_τ][n]
        out = np.empty_like(u)
        out[:, order] = u
        return out

    def __call__(self, idxs, taus, ks=None, ktransform=np.log):
        scalar_i = np.isscalar(idxs) or isinstance(idxs, str)
        scalar_t = np.isscalar(taus)
        scalar_k = ks is not None and np.isscalar(ks)
        idxs = [idxs] if scalar_i else list(idxs)
        taus = np.atleast_1d(np.asarray(taus, dtype=np.float64))
        if ks is None:
            if not all(isinstance(i, str) for i in idxs):
                raise ValueError("perturbation unknowns need wavenumbers: sol(is, τs, ks)")
            out = np.stack([_hermite(self.bg.t, self.bg.u[:, BG_UNKNOWNS[i]], self.bg.du[:, BG_UNKNOWNS[i]], taus)
                            for i in idxs])
        else:
            if self.ks is None or len(self.ks) == 0:
                raise ValueError("No perturbations solved for. Pass ks to solve().")
            if np.any(np.diff(self.ks) < 0):
                raise ValueError("Solution wavenumbers are not sorted in ascending order")
            ks = np.atleast_1d(np.asarray(ks, dtype=np.float64))
            kmin, kmax = self.ks[0], self.ks[-1]
            if ks.min() < kmin:
                raise ValueError(f"Requested wavenumber k = {ks.min()} is below the minimum solved wavenumber {kmin}")
            if ks.max() > kmax:
                raise ValueError(f"Requested wavenumber k = {ks.max()} is above the maximum solved wavenumber {kmax}")
            if not all(isinstance(i, (int, np.integer)) for i in idxs):
                raise ValueError("perturbation unknowns are given by their index in the state vector")
            i2 = np.searchsorted(self.ks, ks, side="left")
            i1 = np.maximum(i2 - 1, 0)
            at = ks == self.ks[np.minimum(i2, len(self.ks) - 1)]
            i1 = np.where(at, i2, i1)                     # exactly on a solved mode
            modes = np.unique(np.concatenate([i1, i2]))
            U = self._modes_at(modes, taus)[:, :, idxs]   # [n_modes][n_τ][n_is]
            pos = {m: j for j, m in enumerate(modes)}
            out = np.zeros((len(idxs), len(taus), len(ks)))
            for ik, k in enumerate(ks):
                v1, v2 = U[pos[i1[ik]]], U[pos[i2[ik]]]
                if i1[ik] == i2[ik]:
                    v = v1
                else:
                    w = (ktransform(k) - ktransform(self.ks[i1[ik]])) / (ktransform(self.ks[i2[ik]]) -
                                                                        ktransform(self.ks[i1[ik]]))
                    v = v1 + (v2 - v1) * w
                out[:, :, ik] = v.T
        if ks is not None and scalar_k:
            out = out[:, :, 0]
        if scalar_t:
            out = out[:, 0] if out.ndim == 2 else out[:, 0, ...]
        if scalar_i:
            out = out[0]
        return out

    def __repr__(self):
        s = f"Cosmology solution for model {self.prob.M.name}\n  Background: return code Success; {len(self.bg.t)} points"
        if self.pts is not None:
            codes = sorted({RETCODES[int(r)] for r in self.pts["retcode"]})
            n = self.pts["stats"][:, 0]
            s += (f"\n  Perturbations: return codes {', '.join(codes)}; solved with Rodas5P (MI355X); "
                  f"{n.min()}-{n.max()} steps; x{len(self.ks)} k ∈ [{self.ks.min()}, {self.ks.max()}] H₀/c")
        return s


def _check_ks(ks):
    ks = np.asarray(ks)
    if ks.dtype.kind not in "fiu":
        raise TypeError(f"wavenumbers must be real Float64 numbers, got dtype {ks.dtype} "
                        "(dual numbers and other element types are not supported by the GPU solver)")
    return np.ascontiguousarray(ks, dtype=np.float64).ravel()


def _ivini(ptivini, ks, t0, t1):
    if callable(ptivini):
        ti = np.array([ptivini(k) for k in ks], dtype=np.float64)
    else:
        ti = np.full(len(ks), float(ptivini))
    return np.clip(ti, t0, t1)  # clamp to the background timespan (solve.jl:527)


def solve(prob, ks=None, bgopts=None, ptopts=None, ptivini=-math.inf, saveat=None, k_unit=None, lanes=0,
          verbose=False, timing=None):
    """Background on the host, perturbation modes on the GPU (one persistent launch)."""
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    bg = solvebg(prob, bgopts["reltol"], bgopts["abstol"])
    if ks is None or len(np.atleast_1d(ks)) == 0:
        return CosmologySolution(prob, bg, None, None)
    ks = _check_ks(ks)
    if k_unit == "1/Mpc":
        ks = C.k_dimensionless(ks, prob.P[P_H])
    elif k_unit is not None:
        raise ValueError("k_unit must be None (H0/c) or '1/Mpc'")
    gm = prob.gpu
    gm.upload_background(0, bg.t, bg.u, bg.du, bg.P)
    t0 = _ivini(ptivini, ks, bg.t[0], bg.t[-1])
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, rtol=ptopts["reltol"], atol=ptopts["abstol"],
                   maxiters=ptopts["maxiters"], saveat=saveat, lanes=lanes)
    ms = plan.run()
    if timing is not None:
        timing["kernel_ms"] = ms
    res = plan.fetch(save=saveat is not None)
    res["pk"] = plan.pk()
    plan.close()
    for i, r in enumerate(res["retcode"]):
        if r != 0:
            warnings.warn(f"Perturbation (mode k = {ks[i]}) solution failed with return code {RETCODES[int(r)]}.")
        elif verbose:
            print(f"\rSolved perturbations for wavenumber k = {ks[i]}", end="")
    return CosmologySolution(prob, bg, ks, res, t0=t0, ptopts=ptopts)


def issuccess(sol):
    return sol.bg.retcode == "Success" and (sol.pts is None or bool(np.all(sol.pts["retcode"] == 0)))


def spectrum_primordial(k, prob):
    """P0(k) = 2π² As (k/k_pivot)^(ns-1) / k³ for dimensionless k (fourier.jl:14-23)."""
    k = np.asarray(k, dtype=np.float64)
    P = prob.P
    return 2 * math.pi**2 * P[P_AS] / k**3 * (k / P[P_KPIVOT]) ** (P[P_NS] - 1)


def spectrum_matter(prob, ks, ktini=1e-2, tauinimax=1e-4, **kw):
    """P(k, τ0) of the total matter overdensity (fourier.jl:79-83 with kτini, τinimax)."""
    sol = solve(prob, ks, ptivini=lambda k: min(ktini / k, tauinimax), **kw)
    return sol.pts["pk"], sol


def species_mask(M, mode):
    """Bit mask of the species of a spectrum_matter mode: "m" (all matter: c, b and massive ν when
    present) or any combination of "c", "b", "h" (fourier.jl:39-52)."""
    bits = {"c": 1, "b": 2, "h": 4}
    if mode == "m":
        return 7 if M.has_h else 3
    mask = 0
    for ch in mode:
        if ch not in bits:
            raise ValueError(f"unknown species '{ch}' in matter power spectrum mode '{mode}'")
        if ch == "h" and not M.has_h:
            raise ValueError(f"mode '{mode}': the model has no massive neutrinos")
        mask |= bits[ch]
    if mask == 0:
        raise ValueError("empty matter power spectrum mode")
    return mask


def spectrum_matter_modes(modes, prob, ks, taus=None, ktini=1e-2, tauinimax=1e-4, bgopts=None, ptopts=None,
                          k_unit=None):
    """spectrum_matter(modes, prob, k[, τ]) (fourier.jl:79-101): P(k, τ) = P0(k) Δ_S(τ)² for each
    mode S ("m", "c", "b", "h", "cb", "cbh", ...).  With taus, the states are saved at those times in
    the same launch and the result is [n_modes][n_τ][n_k]; without, it is [n_modes][n_k] at τ0."""
    ks = _check_ks(ks)
    if k_unit == "1/Mpc":
        ks = C.k_dimensionless(ks, prob.P[P_H])
    elif k_unit is not None:
        raise ValueError("k_unit must be None (H0/c) or '1/Mpc'")
    masks = [species_mask(prob.M, m) for m in modes]
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    bg = solvebg(prob, bgopts["reltol"], bgopts["abstol"])
    gm = prob.gpu
    gm.upload_background(0, bg.t, bg.u, bg.du, bg.P)
    t0 = np.clip(np.minimum(ktini / ks, tauinimax), bg.t[0], bg.t[-1])
    saveat = None if taus is None else np.asarray(taus, dtype=np.float64)
    if saveat is not None and (saveat.min() < bg.t[0] or saveat.max() > bg.t[-1] or np.any(np.diff(saveat) < 0)):
        raise ValueError("τs must be ascending and within the background timespan")
    from .solver import SAVE_STATES
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, rtol=ptopts["reltol"], atol=ptopts["abstol"],
                   maxiters=ptopts["maxiters"], saveat=saveat, save=SAVE_STATES if saveat is not None else 0)
    try:
        plan.run()
        res = plan.fetch()
        for i, r in enumerate(res["retcode"]):
            if r != 0:
                warnings.warn(f"Perturbation (mode k = {ks[i]}) solution failed with return code {RETCODES[int(r)]}.")
        P = plan.pk_modes(masks, at_saves=saveat is not None)
    finally:
        plan.close()
    return P if saveat is not None else P[:, 0, :]


def solve_sets(M, par_sets, ks, bgopts=None, ptopts=None, ktini=1e-2, tauinimax=1e-4, lanes=0, timing=None,
               algo=None):
    """Many parameter sets × the same ks in ONE persistent launch (the batched form of
    parameter_updater + solve, src/solve.jl:272-307 driven in a loop at test/runtests.jl:508-532).
    algo: a kernel variant (symboltz.solver.ALGO_*); None picks the autotuner when the batch holds
    at least twice the one-mode kernel's resident modes (throughput-bound: the variants differ by
    up to 4x there), else the one-mode kernel.  Returns (P[n_sets][n_k], results, backgrounds)."""
    ks = _check_ks(ks)
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    gm = _gpu_model(M)
    probs = [CosmologyProblem(M, pars) for pars in par_sets]
    bgs = solvebg_sets(probs, bgopts["reltol"], bgopts["abstol"], timing=timing)   # all sets at once on the GPU
    inst_set, inst_k, inst_t0 = [], [], []
    for s, bg in enumerate(bgs):
        gm.upload_background(s, bg.t, bg.u, bg.du, bg.P)
        inst_set.append(np.full(len(ks), s, dtype=np.int32))
        inst_k.append(ks)
        inst_t0.append(np.clip(np.minimum(ktini / ks, tauinimax), bg.t[0], bg.t[-1]))
    if algo is None:
        from .autotune import capacity
        from .solver import ALGO_AUTO, ALGO_WAVE
        algo = ALGO_AUTO if (lanes == 0 and len(par_sets) * len(ks) >= 2 * capacity(gm, ALGO_WAVE)) else 0
    plan = gm.plan(np.concatenate(inst_set), np.concatenate(inst_k), np.concatenate(inst_t0),
                   rtol=ptopts["reltol"], atol=ptopts["abstol"], maxiters=ptopts["maxiters"], lanes=lanes, algo=algo)
    ms = plan.run()
    if timing is not None:
        timing["kernel_ms"] = ms
        timing["algo"] = plan.algo
    res = plan.fetch()
    pk = plan.pk().reshape(len(par_sets), len(ks))
    plan.close()
    return pk, res, bgs


_multi_models = {}


def spectrum_matter_problems(probs, ks, bgopts=None, ptopts=None, ktini=1e-2, tauinimax=1e-4, timing=None):
    """P(k, τ0) [n_probs][n_k] of problems of one or several models in ONE persistent launch: the
    problems of each model batched as `solve_sets` does, and the modes of all models in one queue,
    longest first (symboltz.multi).  The batched form of calling `spectrum_matter` for every problem
    of a mixed list.  Per-mode failures warn as `solve` does and give NaN."""
    from .multi import MultiModel
    ks = _check_ks(ks)
    bgopts = dict(reltol=1e-7, abstol=1e-7) | (bgopts or {})
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    models = []
    for q in probs:
        if q.M not in models:
            models.append(q.M)
    if len(models) > 8:
        raise ValueError("one launch holds at most 8 models")
    gms = [_gpu_model(M) for M in models]
    key = tuple(M.key() for M in models)
    if key not in _multi_models:
        _multi_models[key] = MultiModel(gms)
    mm = _multi_models[key]
    plans, where = [], []
    for M, gm in zip(models, gms):
        idx = [i for i, q in enumerate(probs) if q.M == M]
        bgs = solvebg_sets([probs[i] for i in idx], bgopts["reltol"], bgopts["abstol"])
        for s, bg in enumerate(bgs):
            gm.upload_background(s, bg.t, bg.u, bg.du, bg.P)
        inst_set = np.repeat(np.arange(len(idx), dtype=np.int32), len(ks))
        inst_k = np.tile(ks, len(idx))
        inst_t0 = np.concatenate([np.clip(np.minimum(ktini / ks, tauinimax), bg.t[0], bg.t[-1]) for bg in bgs])
        plans.append(gm.plan(inst_set, inst_k, inst_t0, rtol=ptopts["reltol"], atol=ptopts["abstol"],
                             maxiters=ptopts["maxiters"]))
        where.append(idx)
    out = np.full((len(probs), len(ks)), np.nan)
    try:
        ms = mm.run(plans)
        if timing is not None:
            timing["kernel_ms"] = ms
        for plan, idx in zip(plans, where):
            pk = plan.pk().reshape(len(idx), len(ks))
            rc = plan.fetch()["retcode"].reshape(len(idx), len(ks))
            for j, i in enumerate(idx):
                for q in np.nonzero(rc[j])[0]:
                    warnings.warn(f"Perturbation (mode k = {ks[q]}) solution failed with return code "
                                  f"{RETCODES[int(rc[j][q])]}.")
                out[i] = np.where(rc[j] == 0, pk[j], np.nan)
    finally:
        for plan in plans:
            plan.close()
    return out


def spectrum_matter_jacobian(prob, ks, names, relstep=1e-3, ktini=1e-2, tauinimax=1e-4, bgopts=None, ptopts=None,
                             timing=None):
    """∂ log P(k) / ∂ log θ for the parameters `names` (e.g. ["Omega_c0", "Omega_b0"]) at the
    problem's values: returns (J [n_k][n_θ], P(k) at θ).

    Sensitivities by central differences in log θ with the step of the reference's test
    (FiniteDiff's central rule, h = relstep · max(|log θ|, 1); test/runtests.jl:363-377), all
    2 n_θ + 1 cosmologies in ONE persistent launch (`solve_sets`).  The reference differentiates
    this map with ForwardDiff (dual numbers through the solvers); the GPU path takes the
    differences instead, which the reference's test holds equal to atol 1e-3."""
    names = list(names)
    base = dict(prob.pars)
    sets = [base]
    hs = []
    for n in names:
        x = math.log(base[n])
        h = relstep * max(abs(x), 1.0)
        hs.append(h)
        sets.append(dict(base, **{n: math.exp(x + h)}))
        sets.append(dict(base, **{n: math.exp(x - h)}))
    pk, res, _ = solve_sets(prob.M, sets, ks, bgopts=bgopts, ptopts=ptopts, ktini=ktini, tauinimax=tauinimax,
                            timing=timing)
    if not np.all(res["retcode"] == 0):
        raise RuntimeError("a perturbation solve failed while differencing the power spectrum")
    lp = np.log(pk)
    J = np.stack([(lp[1 + 2 * i] - lp[2 + 2 * i]) / (2 * hs[i]) for i in range(len(names))], axis=1)
    return J, pk[0]


def spectrum_matter_tangent(prob, ks, names, hlog=1e-3, ktini=1e-2, tauinimax=1e-4, ptopts=None, bg_tol=1e-10,
                            timing=None):
    """∂ log P(k) / ∂ log θ by forward-mode sensitivities on the GPU: returns (J [n_k][n_θ], P(k)).

    The reference differentiates `log ∘ spectrum_matter ∘ parameter_updater` with ForwardDiff
    (test/runtests.jl:363-377).  Here every (k, θ) instance integrates the perturbations u together
    with s = ∂u/∂log θ in one launch (sbz_plan_create_tangent): the augmented linear system on the
    primal's adaptive step sequence.  The background enters through its θ-dependence at fixed τ,
    the central difference of two background solves at log θ ± hlog with tolerance bg_tol (cheap:
    5 unknowns per set), and the moving end time τ0(θ) is included in the P(k) derivative."""
    names = list(names)
    ks = _check_ks(ks)
    ptopts = dict(reltol=1e-5, abstol=1e-5, maxiters=100000) | (ptopts or {})
    base = dict(prob.pars)
    bg0 = solvebg(prob)
    shifted = []
    for n in names:
        x = math.log(base[n])
        shifted += [dict(base, **{n: math.exp(x + hlog)}), dict(base, **{n: math.exp(x - hlog)})]
    bgs = [solvebg(CosmologyProblem(prob.M, p, prob.ivspan), bg_tol, bg_tol) for p in shifted]
    gm = prob.gpu
    gm.upload_background(0, bg0.t, bg0.u, bg0.du, bg0.P)
    for s_, bg in enumerate(bgs):
        gm.upload_background(1 + s_, bg.t, bg.u, bg.du, bg.P)
    nk, nt = len(ks), len(names)
    t0 = np.clip(np.minimum(ktini / ks, tauinimax), bg0.t[0], bg0.t[-1])
    j = np.repeat(np.arange(nt), nk)
    plan = gm.plan_tangent(np.zeros(nk * nt, np.int32), 1 + 2 * j, 2 + 2 * j, np.full(nk * nt, 0.5 / hlog),
                           np.tile(ks, nt), np.tile(t0, nt), rtol=ptopts["reltol"], atol=ptopts["abstol"],
                           maxiters=ptopts["maxiters"])
    try:
        ms = plan.run()
        if timing is not None:
            timing["kernel_ms"] = ms
        r = plan.fetch()
        pk, d = plan.pk_tangent()
    finally:
        plan.close()
    if not np.all(r["retcode"] == 0):
        raise RuntimeError("a perturbation solve failed in the tangent launch")
    return d.reshape(nt, nk).T, pk[:nk]


def reionization_optical_depth(M, pars, reltol=1e-7, abstol=1e-7):
    """Optical depth of the reionized electrons: κ(τ0) with the tanh reionization terms minus κ(τ0)
    without them (both from full background solves; κ is the reference's _κ, baryons.jl:175-177).

    Not in the reference, which has no τ_reion parameter; it serves parameter sets drawn in
    τ_reion (see `z_re_for_optical_depth`)."""
    from .models import P_REION
    P = make_params(M, pars)
    k_on = _lib.background_solve(P, reltol, abstol)[0][P_KAPPA0]
    P[P_REION] = 0.0
    k_off = _lib.background_solve(P, reltol, abstol)[0][P_KAPPA0]
    return abs(k_on - k_off)


def z_re_for_optical_depth(M, pars, tau_reion, z_lo=3.0, z_hi=30.0, xtol=1e-8, reltol=1e-7, abstol=1e-7):
    """The hydrogen reionization redshift `z_re` whose reionization optical depth is `tau_reion`
    (Brent's method on `reionization_optical_depth`, which grows monotonically with z_re).

    Not in the reference (extension for parameter sets given in τ_reion)."""
    from scipy.optimize import brentq
    from .models import P_REION
    P = make_params(M, dict(pars, z_re=z_lo))
    P[P_REION] = 0.0
    k_off = _lib.background_solve(P, reltol, abstol)[0][P_KAPPA0]

    def g(z):
        Pz = make_params(M, dict(pars, z_re=z))
        return abs(_lib.background_solve(Pz, reltol, abstol)[0][P_KAPPA0] - k_off) - tau_reion

    return brentq(g, z_lo, z_hi, xtol=xtol)
