"""Background stage (the input of the hot path): the host solver of libsbz_hip.so.

Checks mirror test/runtests.jl: initial a ≈ sqrt(Ω_r) τ and τ ≈ 1/ℋ (:240-246), a(τ0) = 1 and
κ(τ0) = 0 (:332-340), τrec at the visibility peak (:63-68); recombination history against the
standard RECFAST numbers; the momentum quadrature against its closed form (:548-559)."""
import math

import numpy as np
import pytest
from scipy.special import zeta

from symboltz import _lib, models


def test_library_loads_and_reports():
    L = _lib.lib()
    assert L.sbz_version() >= 100
    assert _lib.device_count() >= 0


def test_initial_scale_factor_and_conformal_time(fixture5):
    S = fixture5
    P = S.P
    Omr = P[models.P_OMG] + P[models.P_OMNU] + P[models.P_OMH] * np.sum(
        P[models.P_W:models.P_W + 4] * P[models.P_X:models.P_X + 4]) / P[models.P_IRHO0]
    tau = S.t[0]
    assert tau == 1e-6
    assert abs(S.u[0, 0] - math.sqrt(Omr) * tau) < 1e-10
    Hc = S.du[0, 0] / S.u[0, 0]  # ℋ = a'/a
    assert abs(tau - 1.0 / Hc) < 1e-10


def test_today(fixture5):
    S = fixture5
    assert abs(S.u[-1, 0] - 1.0) < 1e-7
    assert S.P[models.P_TAU0] == S.t[-1]
    assert S.P[models.P_KAPPA0] == S.u[-1, 4]
    b = S.O.background(S.t[-1])
    assert b["kappa"] == 0.0
    assert abs(b["a"] - 1.0) < 1e-7


def test_recombination_history(fixture5):
    S = fixture5
    z = 1.0 / S.u[:, 0] - 1.0
    XH = np.interp(1100.0, z[::-1], S.u[::-1, 1])
    assert 0.12 < XH < 0.18            # RECFAST 1.5: x_e(1100) ≈ 0.15
    zrec = 1.0 / np.interp(S.P[models.P_TAUREC], S.t, S.u[:, 0]) - 1.0
    assert 1050 < zrec < 1110          # visibility peak z ≈ 1090
    assert np.all(np.diff(S.t) > 0)


def test_background_rhs_matches_knot_derivative(fixture5):
    S = fixture5
    du = np.zeros(5)
    L = _lib.lib()
    for i in [10, len(S.t) // 2, len(S.t) - 2]:
        u = np.ascontiguousarray(S.u[i])
        L.sbz_background_rhs(_lib.ptr(S.P), _lib.ptr(u), _lib.ptr(du))
        assert abs(du[0] / S.du[i, 0] - 1) < 1e-12     # a' = a ℋ exactly
        assert abs(du[4] / S.du[i, 4] - 1) < 1e-12     # κ' exactly


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8])
def test_momentum_quadrature_closed_form(n):
    x, W = models.quadrature(4)
    num = np.sum(W * x ** (n - 2))
    anal = math.factorial(n) * (1 - 2.0 ** -n) * zeta(n + 1)
    for N in range(1, 6):
        assert abs(num / anal - 1) <= 10.0 ** (-6 + n - N)


def test_momentum_quadrature_exact_polynomials():
    # Gauss rule in u = 1/(1+x/100): exact for polynomials of degree ≤ 2N-1 in u
    from symboltz.quadrature import _discretisation, _x_of_u
    x, W = models.quadrature(4)
    u_nodes = 1.0 / (1.0 + x / 100.0)
    uu, ww = _discretisation(400)
    for p in range(8):
        ref = np.sum(ww * uu**p)
        assert abs(np.sum(W * u_nodes**p) / ref - 1) < 1e-11
