"""Checks of the background and thermodynamics stage that do not rest on the product's own solver
(solvebg, src/solve.jl:427-435; RECFAST in src/models/baryons.jl:14-138):

- a(τ) against the closed form of a radiation + matter universe, a = Ω_m τ²/4 + √Ω_r τ (units
  H0 = c = 1, conformal time), with Ω_γ computed here from the CODATA constants;
- X_H⁺ against the Saha equilibrium of hydrogen before recombination: the RECFAST rate equation
  must track it from slightly above (recombination lags equilibrium) within a few percent in
  1 - X_H⁺ for 1600 ≤ z ≤ 2000;
- the reference's splining test (test/runtests.jl:280-296): every background quantity the
  perturbations recompute from the spline of the default (rtol 1e-7) solve, against the same
  quantities at the knots of a tight (rtol 1e-10) solve, 100 points uniform in log a ∈ [-8, 0],
  with the reference's tolerances;
- the w0wa fluid (dark_energy.jl:24-67) in the oracle: ρ_X(a) against exp(3∫(1+w) d ln a), the
  adiabatic initial condition δ_X = -3/2 (1+w) Ψ with θ_X = θ_c, and P(k) converging to the ΛCDM
  spectrum as w0 → -1 with wa = 0.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from conftest import Setup
from oracle import oracle
from symboltz import _lib, models

# CODATA 2018 and the atomic masses of the reference's constants.jl, restated independently here
KB, HBAR, C, G = 1.380649e-23, 1.054571817e-34, 299792458.0, 6.67430e-11
MPC, ME, AMU, EV = 3.0856775814913673e22, 9.1093837015e-31, 1.66053906660e-27, 1.602176634e-19
MH, MHE = 1.008 * AMU, 4.002602 * AMU
E_H = 13.605693122994 * EV   # hydrogen ionization energy (Rydberg)


def omegas(pars):
    """(Ω_r, Ω_m, H0 in 1/s) of a model with massless neutrinos only."""
    H0 = pars["h"] * 1e5 / MPC
    Omg = math.pi**2 / 15 * (KB * pars["T0"]) ** 4 / (HBAR**3 * C**5) * 8 * math.pi * G / (3 * H0**2)
    Omr = Omg * (1 + pars.get("Neff", 0.0) * 7 / 8 * (4 / 11) ** (4 / 3))
    return Omr, pars["Omega_c0"] + pars["Omega_b0"], H0


def massless_background(pars=None):
    M = models.LCDM(lmax=10, h=False)
    pars = pars or models.parameters_Planck18(M)
    P, t, u, du, _ = _lib.background_solve(models.make_params(M, pars))
    return M, pars, P, t, u, du


def check_scale_factor_closed_form(pars, t, u, du):
    Omr, Omm, _ = omegas(pars)
    m = (u[:, 0] < 0.01) & (t >= 1e-4)
    assert m.sum() > 100
    a_cf = Omm * t[m] ** 2 / 4 + math.sqrt(Omr) * t[m]
    da_cf = Omm * t[m] / 2 + math.sqrt(Omr)
    assert np.max(np.abs(u[m, 0] / a_cf - 1)) < 1e-6
    assert np.max(np.abs(du[m, 0] / da_cf - 1)) < 2e-6


def check_saha(pars, u):
    Omr, Omm, H0 = omegas(pars)
    Y = pars["YHe"]
    fHe = Y / (MHE / MH * (1 - Y))
    rhoc = 3 * H0**2 / (8 * math.pi * G)
    z = 1 / u[:, 0] - 1
    sel = np.flatnonzero((z >= 1600) & (z <= 2000))
    assert len(sel) > 10
    for i in sel:
        T = pars["T0"] * (1 + z[i])      # T_b = T_γ to 1e-6 this early
        nH = (1 - Y) * pars["Omega_b0"] * rhoc * (1 + z[i]) ** 3 / MH
        S = (ME * KB * T / (2 * math.pi * HBAR**2)) ** 1.5 * math.exp(-E_H / (KB * T)) / nH
        b = fHe * u[i, 2]                 # helium electrons (He⁺ from the solve) in n_e
        x = 2 * S / ((b + S) + math.sqrt((b + S) ** 2 + 4 * S))   # x (x + b) / (1 - x) = S
        lag = (1 - u[i, 1]) / (1 - x) - 1
        assert -0.03 < lag <= 0.0, (z[i], u[i, 1], x, lag)


def test_scale_factor_closed_form_radiation_matter():
    M, pars, P, t, u, du = massless_background()
    check_scale_factor_closed_form(pars, t, u, du)


def test_hydrogen_tracks_saha_before_recombination():
    M, pars, P, t, u, du = massless_background()
    check_saha(pars, u)


def test_splined_background_matches_tight_solve():
    # test/runtests.jl:280-296: background quantities as the perturbations see them (spline of
    # the rtol 1e-7 solve, observed quantities recomputed) against the solve itself; here the
    # "solve itself" is a tight rtol 1e-10 solve read at its own knots
    M = models.LCDM(lmax=5)
    pars = models.parameters_Planck18(M)
    P0 = models.make_params(M, pars)
    P, t, u, du, _ = _lib.background_solve(P0.copy())
    Pt, tt, ut, dut, _ = _lib.background_solve(P0.copy(), 1e-10, 1e-10, 100000)
    Od = oracle.Oracle(M, P, t, u, du)
    Ot = oracle.Oracle(M, Pt, tt, ut, dut)
    la = np.log10(ut[:, 0])
    idx = np.unique([int(np.argmin(np.abs(la - x))) for x in np.linspace(-8, 0, 100)])
    idx = idx[(tt[idx] >= t[0]) & (tt[idx] <= t[-1])]
    assert len(idx) > 80

    def vals(O, tau):
        b = O.background(tau)
        v, vd, _, _ = O.visibility(tau)
        return dict(a=b["a"], kd=b["kappadot"], kappa=b["kappa"], v=v, vd=vd, cs2=b["cs2"], T=b["Tb"], Xe=b["Xe"])

    # the reference's tolerances, except T: the reference compares the spline with the dense output
    # of the same solve, this test with a tighter solve, so the default solve's global error enters;
    # T_b = T_γ + ΔT with the unknown ΔT ≈ -T_γ at late times, where ~1e-4 K on ΔT is 2e-5 of T_b
    tol = dict(a=(1e-6, 0), kd=(0, 1e-2), kappa=(0, 1e-4), v=(1e-3, 0), vd=(0, 1e1), cs2=(1e-4, 0), T=(0, 3e-5),
               Xe=(1e-5, 0))
    for i in idx:
        v1, v2 = vals(Ot, tt[i]), vals(Od, tt[i])
        for key, (atol, rtol) in tol.items():
            assert abs(v1[key] - v2[key]) <= atol + rtol * max(abs(v1[key]), abs(v2[key])), (key, tt[i], v1[key], v2[key])


def test_w0wa_density_is_the_integral_of_the_equation_of_state():
    M = models.w0waCDM(lmax=6)
    for w0, wa in ((-0.9, 0.2), (-1.1, -0.4), (-0.8, 0.5)):
        S = Setup(M, dict(models.parameters_Planck18(M), w0=w0, wa=wa))
        rho0 = S.O.background(S.t[-1])["rho_x"]
        for tau in (0.05, 0.5, 1.5, 3.0):
            b = S.O.background(tau)
            a = b["a"]
            w = lambda x: w0 + wa * (1 - x)
            assert abs(b["w_x"] - w(a)) < 1e-14
            # ρ_X(a)/ρ_X(1) = exp(3 ∫_a^1 (1 + w(a')) da'/a')
            I = quad(lambda x: (1 + w(x)) / x, a, 1.0, epsabs=0, epsrel=1e-13)[0]
            assert abs(b["rho_x"] / rho0 / math.exp(3 * I) - 1) < 1e-11, (w0, wa, tau)


def test_w0wa_adiabatic_initial_conditions():
    M = models.w0waCDM(lmax=6)
    S = Setup(M, dict(models.parameters_Planck18(M), w0=-0.9, wa=0.2))
    L = S.O.layout()
    tau = S.t[0]
    b = S.O.background(tau)
    fnu = (b["rho_nu"] + b["rho_h"]) / (b["rho_g"] + b["rho_nu"] + b["rho_h"])
    Psi = 10.0 / (15.0 + 4.0 * fnu)        # Ψ = 20C/(15 + 4 f_ν), C = 1/2 (cosmologies.jl:55,75)
    for k in (1.0, 100.0):
        y = S.O.initial_state(k, tau)
        assert abs(y[L["X"]] / (-1.5 * (1 + b["w_x"]) * Psi) - 1) < 1e-12
        assert abs(y[L["X"] + 1] / y[L["C"] + 1] - 1) < 1e-12
        assert abs(y[L["C"]] / (-1.5 * Psi) - 1) < 1e-12


def test_w0wa_power_spectrum_tends_to_lcdm():
    # with wa = 0 and w0 = -1 + ε the fluid's background is Λ's to O(ε) and its perturbations are
    # O(ε): P(k) must approach the ΛCDM spectrum of the same species content linearly in ε
    Mx, Ml = models.w0waCDM(lmax=6), models.LCDM(lmax=6)
    pars = models.parameters_Planck18(Ml)
    ks = 10 ** np.linspace(-1, 3, 9)
    Pl, rl = Setup(Ml, pars).O.spectrum_matter(ks)
    assert np.all(rl["retcode"] == 0)
    errs = []
    for eps in (1e-2, 1e-3):
        Px, rx = Setup(Mx, dict(pars, w0=-1 + eps, wa=0.0, cs2=1.0)).O.spectrum_matter(ks)
        assert np.all(rx["retcode"] == 0)
        errs.append(np.max(np.abs(Px / Pl - 1)))
    assert errs[1] < 2e-3
    assert 5 < errs[0] / errs[1] < 20, errs


@pytest.mark.gpu
def test_gpu_background_batch_closed_forms():
    """The batched GPU background solve (one thread per set) against the same closed forms: a(τ)
    and the Saha limit, for Latin-hypercube sets of the massless-ν model."""
    from scipy.stats import qmc
    M = models.LCDM(lmax=10, h=False)
    base = models.parameters_Planck18(M)
    X = qmc.scale(qmc.LatinHypercube(d=5, seed=7).random(16), [0.60, 0.020, 0.10, 2.9, 0.92],
                  [0.75, 0.024, 0.14, 3.2, 1.00])
    sets = [dict(base, h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=la, ns=ns) for h, wb, wc, la, ns in X]
    Ps = np.stack([models.make_params(M, p) for p in sets])
    P, knots, nrej, status, ms = _lib.background_solve_batch(Ps)
    assert np.all(status == 0)
    for p, (t, u, du) in zip(sets, knots):
        check_scale_factor_closed_form(p, t, u, du)
        check_saha(p, u)
