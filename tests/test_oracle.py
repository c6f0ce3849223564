"""Pin the CPU oracle (oracle/sbz_oracle.cpp) before trusting it.

- golden vectors of the reference: test/class_Pk.dat (CLASS P(k) for the reference's fixture
  cosmology), compared as the reference does, with isapprox's norm-wise rtol 1e-3
  (test/runtests.jl:872-876), and element-wise on 1 ≤ k ≤ 1000 H0/c;
- the reference's self-convergence gate: P(k) at rtol 1e-5 vs rtol 1e-10 with kτini = 0 within
  1e-3 (test/runtests.jl:624-630);
- initial-condition invariants of test/runtests.jl:232-278;
- mathematics: f linear in u, ∂f/∂τ against a finite difference in τ, success/failure semantics
  (k = 0 and k = NaN fail, test/runtests.jl:31,360).
"""
import numpy as np
import pytest

from conftest import GOLDEN, Setup


def test_class_pk_fixture(fixture5):
    S = fixture5
    d = np.loadtxt(GOLDEN / "class_Pk.dat")
    ks, Pc = d[:, 0], d[:, 1]
    P, r = S.O.spectrum_matter(ks)
    assert np.all(r["retcode"] == 0)
    normwise = np.linalg.norm(P - Pc) / max(np.linalg.norm(P), np.linalg.norm(Pc))
    assert normwise < 1e-3
    mid = (ks >= 1) & (ks <= 1000)
    assert np.max(np.abs(P[mid] / Pc[mid] - 1)) < 1e-3


def test_self_convergence(fixture5):
    S = fixture5
    ks = 10 ** np.linspace(-1, 4, 40)
    P1, r1 = S.O.spectrum_matter(ks, kt_ini=0.0, tau_ini_max=0.0)
    # the rtol = 1e-10 run needs ~1e5 steps at k = 1e4, above the default maxiters
    P0, r0 = S.O.spectrum_matter(ks, kt_ini=0.0, tau_ini_max=0.0, rtol=1e-10, atol=1e-10, maxiters=10**6)
    assert np.all(r0["retcode"] == 0) and np.all(r1["retcode"] == 0)
    assert np.max(np.abs(P1 / P0 - 1)) < 1e-3


def test_initial_conditions(fixture5):
    S, O = fixture5, fixture5.O
    L = O.layout()
    tau = S.t[0]
    ks = np.array([1e-1, 1e0]) / 4474.0 * 4474.0  # two k in H0/c units
    ks = np.array([0.1, 1.0]) * (2997.92458 / 0.6736)  # 0.1/Mpc and 1/Mpc
    u = [O.initial_state(k, tau) for k in ks]
    # F_l(τ_ini) ∝ k^l for l = 0..3
    for l in range(4):
        r = u[0][L["F"] + l] / u[1][L["F"] + l]
        assert abs(r / (ks[0] / ks[1]) ** l - 1) < 1e-8
    # Φ/Ψ = 1 + 2/5 fν
    b = O.background(tau)
    fnu = (b["rho_nu"] + b["rho_h"]) / (b["rho_g"] + b["rho_nu"] + b["rho_h"])
    Psi = 10.0 / (15.0 + 4.0 * fnu)
    for y in u:
        assert abs(y[L["Phi"]] / Psi - (1 + 0.4 * fnu)) < 1e-4
    # δ/(1+w) and u/k = θ/k² agree across species (rtol 1e-3)
    for k, y in zip(ks, u):
        d_over = [y[L["C"]], y[L["B"]], y[L["F"]] / (4 / 3), y[L["N"]] / (4 / 3)]
        th = [y[L["C"] + 1] / k**2, y[L["B"] + 1] / k**2, 3 * k * y[L["F"] + 1] / 4 / k**2,
              3 * k * y[L["N"] + 1] / 4 / k**2]
        assert (max(d_over) - min(d_over)) / abs(np.mean(d_over)) < 1e-3
        assert (max(th) - min(th)) / abs(np.mean(th)) < 1e-3


def test_linearity_and_time_derivative(fixture5):
    S, O = fixture5, fixture5.O
    rng = np.random.default_rng(0)
    n = O.n
    for tau in [1e-4, 0.02, 0.06, 1.0, 3.0]:
        for k in [0.5, 30.0, 2000.0]:
            x, y = rng.standard_normal(n), rng.standard_normal(n)
            a, b = 0.7, -1.3
            lhs = O.rhs(k, tau, a * x + b * y)
            rhs = a * O.rhs(k, tau, x) + b * O.rhs(k, tau, y)
            assert np.allclose(lhs, rhs, rtol=1e-10, atol=1e-10 * np.abs(lhs).max())
            J = O.jacobian(k, tau)
            assert np.allclose(J @ x, O.rhs(k, tau, x), rtol=1e-10, atol=1e-10 * np.abs(J @ x).max())
            h = tau * 1e-6
            fd = (O.rhs(k, tau + h, x) - O.rhs(k, tau - h, x)) / (2 * h)
            dt = O.dfdt(k, tau, x)
            scale = np.abs(fd).max()
            assert np.max(np.abs(dt - fd)) < 1e-4 * scale


def test_failure_retcodes(fixture5):
    S, O = fixture5, fixture5.O
    r = O.solve(np.array([1.0, 0.0, np.nan]), np.array([1e-4, 1e-4, 1e-4]), nthreads=1)
    assert r["retcode"][0] == 0
    assert r["retcode"][1] != 0 and r["retcode"][2] != 0


def test_visibility_derivatives_by_finite_differences(fixture5):
    """The sources' v, v', v'' (second-order Taylor arithmetic through the background, baryons.jl:179)
    against 5-point central differences of v(τ) = -κ̇ e^{-κ} inside one spline interval (where the
    cubic Hermite background is smooth), across recombination, reionization and late times."""
    S, O = fixture5, fixture5.O
    n = O.n
    L = O.layout()
    ti = S.t
    taurec = S.P[25]
    picks = [np.searchsorted(ti, x) for x in (0.8 * taurec, taurec, 1.2 * taurec, 0.5 * ti[-1], 0.9 * ti[-1])]
    vis = lambda t: -O.background(t)["kappadot"] * np.exp(-O.background(t)["kappa"])
    for i in picks:
        tau = 0.5 * (ti[i] + ti[i + 1])
        v, vd, vdd, eK = O.visibility(tau)
        assert abs(v / vis(tau) - 1) < 1e-13
        assert abs(eK - np.exp(-O.background(tau)["kappa"])) < 1e-15
        # v' against differences of v, v'' against differences of v' (one derivative per stencil
        # keeps the rounding of the differences below 1e-6)
        h = 0.1 * (ti[i + 1] - ti[i])
        f = {j: vis(tau + j * h) for j in (-2, -1, 1, 2)}
        vd_fd = (f[-2] - 8 * f[-1] + 8 * f[1] - f[2]) / (12 * h)
        assert abs(vd - vd_fd) <= 3e-6 * abs(vd), (i, vd, vd_fd)
        h = 0.05 * (ti[i + 1] - ti[i])
        g = {j: O.visibility(tau + j * h)[1] for j in (-2, -1, 1, 2)}
        vdd_fd = (g[-2] - 8 * g[-1] + 8 * g[1] - g[2]) / (12 * h)
        assert abs(vdd - vdd_fd) <= 1e-5 * abs(vdd), (i, vdd, vdd_fd)
        # S_E = 3 v Π_γ / (16 (k χ)²) of a state with only Π_γ = F2 = 1 reads the same v
        u = np.zeros(n)
        u[L["F"] + 2] = 1.0
        k = 1.0
        chi = S.t[-1] - tau
        SE = O.sources(k, tau, u)[1]
        assert abs(SE / (3 / 16 * v / (k * chi) ** 2) - 1) < 1e-10


def test_sparse_path_matches_dense(massless10):
    """The oracle's sparse path (J by column coloring of the linear RHS, static-pattern LU in a
    minimum-degree order without pivoting: the CPU baseline's solver) against its dense path
    (J by unit vectors, partial-pivot LU): same states to rounding, same step counts (±1), and
    the same CMB sources for k ≤ 3000 H0/c."""
    from symboltz import workloads
    for S in (massless10, Setup(*workloads.massive_nu100())):
        ks = 10 ** np.linspace(-1, 4, 12)
        t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
        saveat = np.linspace(0.05, S.t[-1], 7)
        d = S.O.solve(ks, t0, saveat=saveat, sources=True)
        s = S.O.solve(ks, t0, saveat=saveat, sources=True, sparse=True)
        assert np.all(d["retcode"] == 0) and np.all(s["retcode"] == 0)
        sc = np.abs(d["u"]).max(axis=1, keepdims=True)
        assert np.max(np.abs(s["u"] - d["u"]) / sc) < 1e-6
        assert np.max(np.abs(s["stats"][:, 0] - d["stats"][:, 0])) <= 1
        lo = ks <= 3000.0   # above it the late-time ISW term amplifies rounding-level state differences
        for c in range(3):
            scs = np.abs(d["src"][lo, :, c]).max() + 1e-300
            assert np.max(np.abs(s["src"][lo, :, c] - d["src"][lo, :, c])) <= 1e-5 * scs
