"""Line-of-sight integration and the Bessel cache (src/observables/angular.jl:9-48,109-185)."""
import math

import numpy as np
import pytest
from scipy.special import sici, spherical_jn

from oracle import oracle_cmb as oc
from oracle.oracle_los import hermite, los_integrate as los_oracle


def test_oracle_bessel_table_matches_scipy():
    # the recurrence table against scipy's spherical_jn at grid points, including the turning
    # point x ≈ l where Miller's recurrence takes over from the upward one
    ls = np.array([0, 1, 2, 3, 10, 50, 200, 700, 1500, 2500])
    dx, y, dy = oc.bessel_table(ls)
    assert y.shape == (int(20 * 2500 / (2 * math.pi / 15)) + 1, len(ls))
    rng = np.random.default_rng(0)
    for il, l in enumerate(ls):
        i0 = int(l / dx)
        idx = np.unique(np.concatenate([rng.integers(0, y.shape[0] - 1, 200),
                                        np.arange(max(0, i0 - 150), min(y.shape[0] - 1, i0 + 150))]))
        x = idx * dx
        assert np.max(np.abs(y[idx, il] - spherical_jn(l, x))) < 1e-14
        assert np.max(np.abs(dy[idx, il] - spherical_jn(l, x, derivative=True))) < 1e-14
    # padding row duplicates the last grid point
    assert np.array_equal(y[-1], y[-2]) and np.array_equal(dy[-1], dy[-2])


def test_bessel_hermite_interpolation_error():
    ls = np.array([2, 10, 50, 200, 700])
    table = oc.bessel_table(ls)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, table[0] * (table[1].shape[0] - 2), 4000)
    for il, l in enumerate(ls):
        # cubic Hermite on dx = 2π/15: error ≲ dx⁴/384 max|j''''|
        assert np.max(np.abs(hermite(table, il, x) - spherical_jn(l, x))) < 2e-5


def test_los_oracle_closed_form():
    # S = 1, l = 0: I_0(k) = ∫_0^χmax j0(kχ) dχ = Si(k χmax)/k up to the trapezoid error
    ls = np.array([0, 1, 5])
    taus = np.linspace(0.0, 3.0, 30001)
    ks = np.array([2.0, 20.0, 200.0])
    table = oc.bessel_table(ls, xmax=ks[-1] * 3.0 * 1.01, dx=2 * math.pi / 600)
    S = np.ones((len(taus), len(ks)))
    I = los_oracle(S, taus, ks, table)
    for ik, k in enumerate(ks):
        # trapezoid error ~ (k Δτ)²/12 of the integral; a wrong weight or χ would be O(1)
        assert abs(I[ik, 0] / (sici(k * 3.0)[0] / k) - 1) < 1e-4
        # l = 1: ∫ j1(x) dx = 1 - j0(x)
        assert abs(I[ik, 1] / ((1 - np.sinc(k * 3.0 / np.pi)) / k) - 1) < 1e-4


def test_cmb_oracle_los_matches_numpy_restatement():
    # the C++ LOS of the C_l oracle (no Limber) against the numpy restatement above
    rng = np.random.default_rng(3)
    ls = np.array([2, 5, 11, 40, 90])
    taus = np.sort(np.concatenate([[0.01], rng.uniform(0.01, 3.0, 58), [3.0]]))
    ks = np.sort(rng.uniform(1.0, 60.0, 7))
    table = oc.bessel_table(ls, xmax=max(20 * ls[-1], ks[-1] * 3.0 * 1.01))
    S = rng.standard_normal((len(taus), len(ks)))
    ref = los_oracle(S, taus, ks, table)
    out = oc.los(S.T[:, :, None], taus, ks, ls, table, [oc.INT_MAX])
    assert np.max(np.abs(out[0] - ref)) <= 1e-13 * np.abs(ref).max()


def test_limber_limit_of_the_oracle():
    # for a source that varies slowly on the scale 1/k, the full LOS integral approaches the
    # Limber value √(π/(2l+1)) S(τ0 - (l+½)/k) / k (angular.jl:139-140)
    ls = np.array([200, 400])
    taus = np.linspace(0.0, 3.0, 60001)
    S1 = np.exp(-((taus - 1.5) / 0.4) ** 2)
    ks = np.array([300.0, 500.0])
    table = oc.bessel_table(ls, xmax=ks[-1] * 3.0 * 1.01, dx=2 * math.pi / 60)
    S = np.repeat(S1[None, :, None], len(ks), axis=0)
    full = oc.los(S, taus, ks, ls, table, [oc.INT_MAX])[0]
    lim = oc.los(S, taus, ks, ls, table, [0])[0]
    assert np.max(np.abs(lim / full - 1)) < 2e-2
    # the Limber branch itself equals the closed form at the sampled point
    for ik, k in enumerate(ks):
        for il, l in enumerate(ls):
            chi = (l + 0.5) / k
            exact = math.sqrt(math.pi / (2 * l + 1)) * math.exp(-((3.0 - chi - 1.5) / 0.4) ** 2) / k
            assert abs(lim[ik, il] / exact - 1) < 1e-6


@pytest.mark.gpu
def test_gpu_bessel_table_matches_oracle():
    from symboltz.bessel import SphericalBesselCache
    print("case: Bessel table l ≤ 2500 on the GPU", flush=True)
    ls = np.array([0, 1, 2, 3, 7, 30, 100, 400, 1000, 2000, 2500])
    jl = SphericalBesselCache(ls)
    dx, y, dy = oc.bessel_table(ls)
    assert abs(jl.dx - dx) == 0.0 and jl.y.shape == y.shape
    assert np.max(np.abs(jl.y - y)) < 1e-14
    assert np.max(np.abs(jl.dy - dy)) < 1e-14
