"""The emitted model code (host build of the same source the GPU runs) against the oracle.

RHS, the W = I/(γh) - J values on the static LU pattern, the LU + triangular solves, ∂f/∂τ,
the initial conditions and Δ_m are compared with the oracle's independent restatement on the
same background at many (k, τ) points."""
import numpy as np
import pytest

TAUS = [1e-6, 3e-5, 1e-3, 0.02, 0.06, 0.3, 1.5, 3.0]
KS = [0.05, 1.0, 40.0, 900.0, 3e4]


@pytest.fixture(params=["fixture5", "massless10"])
def pair(request, probe5, probe10):
    S = request.getfixturevalue(request.param)
    return S, (probe5 if request.param == "fixture5" else probe10)


def test_rhs(pair):
    S, pr = pair
    rng = np.random.default_rng(1)
    for tau in TAUS:
        for k in KS:
            u = rng.standard_normal(pr.n)
            a, b = pr.f(S, k, tau, u), S.O.rhs(k, tau, u)
            assert np.max(np.abs(a - b)) <= 1e-11 * np.abs(b).max(), (tau, k)


def test_w_matrix_and_lu_solve(pair):
    S, pr = pair
    rng = np.random.default_rng(2)
    for tau in TAUS:
        for k in KS:
            J = S.O.jacobian(k, tau)
            for h in [1e-7, 1e-3, 0.1]:
                ig = 1.0 / (0.21193756319429014 * h)
                Wd = np.eye(pr.n) * ig - J
                b = rng.standard_normal(pr.n)
                ok, x = pr.lu_solve(pr.w(S, k, tau, ig), b)
                assert ok
                xr = np.linalg.solve(Wd, b)
                res = np.abs(Wd @ x - b).max() / (np.abs(Wd).max() * np.abs(x).max())
                assert res < 1e-13, (tau, k, h, res)
                # forward error of two backward-stable solvers differs by at most ~cond(W)·ε
                cond = np.linalg.cond(Wd)
                assert np.max(np.abs(x - xr)) <= 1e-14 * cond * np.abs(xr).max() + 1e-300


def test_time_derivative(pair):
    S, pr = pair
    rng = np.random.default_rng(3)
    for tau in TAUS:
        for k in KS:
            u = rng.standard_normal(pr.n)
            a, b = pr.dfdt(S, k, tau, u), S.O.dfdt(k, tau, u)
            assert np.max(np.abs(a - b)) <= 1e-9 * np.abs(b).max(), (tau, k)


def test_initial_conditions_and_overdensity(pair):
    S, pr = pair
    for k in KS:
        for tau in [1e-6, 1e-4]:
            a, b = pr.ic(S, k, tau), S.O.initial_state(k, tau)
            assert np.allclose(a, b, rtol=1e-12, atol=1e-300)
            assert abs(pr.dm(S, k, 3.0, b) / S.O.delta_m(k, 3.0, b) - 1) < 1e-12


def test_cmb_sources(pair):
    S, pr = pair
    rng = np.random.default_rng(4)
    taurec = S.P[25]
    for tau in [0.01, 0.05, taurec, 0.2, 1.0, 3.0]:
        for k in KS:
            u = rng.standard_normal(pr.n)
            a, b = pr.src(S, k, tau, u), S.O.sources(k, tau, u)
            scale = np.abs(b).max() + 1e-300
            assert np.max(np.abs(a - b)) <= 1e-9 * scale, (tau, k, a, b)
