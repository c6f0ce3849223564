"""Synthetic workloads of the perturbation stage at the sizes the benchmark and the tests use.

Each returns (model, parameter sets, k in 1/Mpc, description).  The cosmology of the single-set
workloads is h = 0.67, ω_b = 0.0224, ω_c = 0.120, T_CMB = 2.7255 K, Y_He = 0.245,
A_s = 2.1e-9, n_s = 0.966, hydrogen reionization at z = 7.7.

- "massless10": ΛCDM with massless ν (N_eff = 3.046), every hierarchy truncated at lmax = 10
  (n = 38 unknowns); 100 k log-spaced in [1e-4, 1]/Mpc.
- "massive_nu100": the same cosmology with one massive ν (0.06 eV, 4 momentum bins, lmax 10),
  N_eff(massless) = 2.046, photon and massless-ν lmax 16 (n = 100); 1024 k log-spaced in
  [1e-5, 10]/Mpc, with the CMB sources on a 2000-point cosgrid of conformal time.
- "lhs128": the n = 38 model over 128 Latin-hypercube parameter sets (seed 0) in h ∈ [0.60, 0.75],
  ω_b ∈ [0.020, 0.024], ω_c ∈ [0.10, 0.14], ln(10¹⁰A_s) ∈ [2.9, 3.2], n_s ∈ [0.92, 1.00],
  τ_reion ∈ [0.04, 0.08] (met by a root-find on z_re), × 512 k log-spaced in [1e-5, 10]/Mpc:
  65,536 instances in one launch.

The reference has no τ_reion parameter (reionization is fixed at z = 7.6711 in baryons.jl:202-203);
z_re is this build's parameter, and τ_reion is converted by `problem.z_re_for_optical_depth`.

- "massive_nu100_lhs64": the n = 100 model over the first 64 "lhs128" parameter sets (their z_re)
  × the 1024 k of "massive_nu100", P(k) only: the same model at a throughput-bound size.
- "mixed": the mixed-model stream (`mixed_families`, `mixed_requests`): requests spread over five
  model families with 38 to 236 unknowns (ΛCDM with massless ν; three degenerate massive ν, the
  surrogate's family; the w0wa fluid; one massive ν at lmax 16; three massive ν at lmax 32), each
  with Latin-hypercube parameters and one k grid.
"""
import math

import numpy as np

BASE = {"h": 0.67, "Omega_b0": 0.0224 / 0.67**2, "Omega_c0": 0.120 / 0.67**2, "T0": 2.7255, "YHe": 0.245,
        "ln_As1e10": math.log(2.1e-9 * 1e10), "ns": 0.966, "z_re": 7.7}

LHS_LO = np.array([0.60, 0.020, 0.10, 2.9, 0.92, 0.04])
LHS_HI = np.array([0.75, 0.024, 0.14, 3.2, 1.00, 0.08])


def lhs_sets(n_sets=128, seed=0, take=None, threads=16):
    """The first `take` (default all) of `n_sets` Latin-hypercube parameter sets over h, ω_b, ω_c,
    ln 10¹⁰A_s, n_s and τ_reion, for the massless-ν lmax-10 model; returns (model, sets)."""
    from concurrent.futures import ThreadPoolExecutor
    from scipy.stats import qmc
    from . import models
    from .problem import z_re_for_optical_depth
    X = qmc.scale(qmc.LatinHypercube(d=6, seed=seed).random(n_sets), LHS_LO, LHS_HI)
    if take is not None:
        X = X[:take]
    M = models.LCDM(lmax=10, h=False)
    sets = [dict(BASE, h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=lnAs, ns=ns, Neff=3.046)
            for h, wb, wc, lnAs, ns, _ in X]
    with ThreadPoolExecutor(max_workers=threads) as ex:   # the background solves release the GIL
        zs = list(ex.map(lambda a: z_re_for_optical_depth(M, a[0], a[1], xtol=1e-6), zip(sets, X[:, 5])))
    for p, z in zip(sets, zs):
        p["z_re"] = z
    return M, sets


def massive_nu100():
    """(model, parameters) of the n = 100 massive-ν model."""
    from . import models
    M = models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)
    return M, dict(BASE, Neff=2.046, N=1, m_eV=0.06)


def source_taus(t_bg0, tau0, n=2000):
    """cosgrid(max(τ_bg0, 0.01), τ0; length = n) (src/utils.jl:285, angular.jl:280-286)."""
    a0, b0 = max(t_bg0, 1e-2), tau0
    x = np.linspace(0.0, 0.5, n)
    cospi = np.where(x == 0.5, 0.0, np.cos(np.pi * x))   # cospi(1/2) = 0 exactly, as Julia's cospi
    return a0 + (b0 - a0) * (1.0 - cospi)


def workload(name):
    """(model, parameter sets, k in 1/Mpc, description) of a named workload."""
    from . import models
    if name == "lhs128":
        M, sets = lhs_sets(128)
        ks = np.logspace(-5, 1, 512)
        desc = ("LCDM massless nu lmax 10, n=38; 128 Latin-hypercube parameter sets (seed 0: h, omega_b, "
                "omega_c, ln 1e10 As, ns, tau_reion via a z_re root-find) x 512 k log-spaced in [1e-5, 10]/Mpc "
                "= 65536 instances in one persistent launch; rtol=atol=1e-5; P(k) at z=0")
        return M, sets, ks, desc
    if name == "massive_nu100":
        M, pars = massive_nu100()
        ks = np.logspace(-5, 1, 1024)
        desc = ("LCDM + 1 massive nu (0.06 eV, 4 bins, lmax 10), photon/massless-nu lmax 16, n=100; "
                "1024 k log-spaced in [1e-5, 10]/Mpc; rtol=atol=1e-5; P(k) at z=0 and CMB sources "
                "(S_T, S_E, S_psi) on a 2000-point cosgrid of conformal time")
        return M, [pars], ks, desc
    if name == "massive_nu100_lhs64":
        M, base = massive_nu100()
        _, lhs = lhs_sets(128, seed=0, take=64)
        sets = [dict(base, **{q: p[q] for q in ("h", "Omega_b0", "Omega_c0", "ln_As1e10", "ns", "z_re")}) for p in lhs]
        ks = np.logspace(-5, 1, 1024)
        desc = ("the n=100 massive-nu model (one massive nu 0.06 eV, 4 bins, lmax 10; photon/massless-nu lmax 16) "
                "over the first 64 Latin-hypercube sets of lhs128 (their h, omega_b, omega_c, ln 1e10 As, ns, z_re) "
                "x 1024 k log-spaced in [1e-5, 10]/Mpc = 65536 instances in one launch; rtol=atol=1e-5; P(k) at z=0")
        return M, sets, ks, desc
    if name == "massless10":
        M = models.LCDM(lmax=10, h=False)
        ks = np.logspace(-4, 0, 100)
        desc = "LCDM massless nu lmax 10, n=38; 100 k log-spaced in [1e-4, 1]/Mpc; rtol=atol=1e-5"
        return M, [dict(BASE, Neff=3.046)], ks, desc
    raise ValueError(f"unknown workload {name}")


def mixed_families():
    """The model families of the mixed-model stream, by state size: ΛCDM with massless ν (lmax 10,
    n = 38), ΛCDM with three degenerate massive ν (lmax 6, n = 54; the surrogate's family), w0waCDM
    (lmax 6, n = 56), ΛCDM with one massive ν (photon/massless-ν lmax 16, lmax 10 per bin, n = 100)
    and ΛCDM at lmax 32 (n = 236)."""
    from . import models
    return [models.LCDM(lmax=10, h=False), models.LCDM(lmax=6), models.w0waCDM(lmax=6),
            models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4), models.LCDM(lmax=32)]


MIXED_SURROGATE_FAMILY = 1


def mixed_requests(n, seed=0, out_of_box=0.0):
    """n requests (family, parameters, surrogate input or None), families in turn.  Parameters from
    one Latin hypercube over h, ω_b, ω_c, ln 10¹⁰A_s, n_s, z_re, Σm_ν ∈ [0.06, 0.30] eV, w0 ∈ [-1.2,
    -0.8], wa ∈ [-0.5, 0.5]; w0wa draws whose w(a) = w0 + wa (1 - a) would cross -1 take -wa (the
    fluid equations divide by 1 + w, dark_energy.jl:55-57).  A fraction `out_of_box` of the
    surrogate family's requests get Σm_ν = 0.45 eV, outside the surrogate's box."""
    from scipy.stats import qmc
    from . import models
    fams = mixed_families()
    lo = np.array([0.60, 0.020, 0.10, 2.9, 0.92, 6.2, 0.06, -1.2, -0.5])
    hi = np.array([0.75, 0.024, 0.14, 3.2, 1.00, 10.3, 0.30, -0.8, 0.5])
    X = qmc.scale(qmc.LatinHypercube(d=len(lo), seed=seed).random(n), lo, hi)
    rng = np.random.default_rng(seed)
    out = []
    for i, (h, wb, wc, la, ns, zre, mnu, w0, wa) in enumerate(X):
        f = i % len(fams)
        M = fams[f]
        p = dict(models.parameters_Planck18(M), h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=la, ns=ns,
                 z_re=zre)
        x = None
        if M.de == "w0wa":
            if (w0 + 1.0) * (w0 + wa + 1.0) < 0:
                wa = -wa
            p.update(w0=w0, wa=wa)
        if M.has_h:
            if f == MIXED_SURROGATE_FAMILY:
                if rng.random() < out_of_box:
                    mnu = 0.45
                p.update(N=3, m_eV=mnu / 3)
                x = np.array([h, wb, wc, la, ns, zre, mnu])
            elif M.lmax_h == 10 and M.lmax_g == 16:
                p.update(N=1, m_eV=mnu, Neff=2.046)
            else:
                p.update(N=3, m_eV=mnu / 3)
        out.append((f, p, x))
    return out
