"""Line-of-sight integration (src/observables/angular.jl:9-48,109-185)."""
import math

import numpy as np
import pytest
from scipy.special import sici, spherical_jn

from symboltz.bessel import SphericalBesselCache
from oracle.oracle_los import los_integrate as los_oracle


def test_bessel_cache_matches_direct_evaluation():
    ls = np.array([2, 10, 50, 200, 700])
    jl = SphericalBesselCache(ls)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, jl.x[-1], 4000)
    for il, l in enumerate(ls):
        d = spherical_jn(l, x)
        # cubic Hermite on dx = 2π/15: error ≲ dx⁴/384 max|j''''|
        assert np.max(np.abs(jl(il, x) - d)) < 2e-5
        # knots are exact
        assert np.allclose(jl(il, jl.x[:-1]), spherical_jn(l, jl.x[:-1]), atol=1e-14)


def test_los_oracle_closed_form():
    # S = 1, l = 0: I_0(k) = ∫_0^χmax j0(kχ) dχ = Si(k χmax)/k up to the trapezoid error
    ls = np.array([0, 1, 5])
    taus = np.linspace(0.0, 3.0, 30001)
    ks = np.array([2.0, 20.0, 200.0])
    jl = SphericalBesselCache(ls, xmax=ks[-1] * 3.0 * 1.01, dx=2 * math.pi / 600)
    S = np.ones((len(taus), len(ks)))
    I = los_oracle(S, taus, ks, jl)
    for ik, k in enumerate(ks):
        # trapezoid error ~ (k Δτ)²/12 of the integral; a wrong weight or χ would be O(1)
        assert abs(I[ik, 0] / (sici(k * 3.0)[0] / k) - 1) < 1e-4
        # l = 1: ∫ j1(x) dx = 1 - j0(x)
        assert abs(I[ik, 1] / ((1 - np.sinc(k * 3.0 / np.pi)) / k) - 1) < 1e-4


@pytest.mark.gpu
def test_los_gpu_matches_oracle():
    from symboltz.los import los_integrate
    print("case: LOS 300 τ x 64 k x 40 l", flush=True)
    rng = np.random.default_rng(1)
    ls = np.unique(np.geomspace(2, 1500, 40).astype(int))
    taus = np.sort(np.concatenate([[0.01], rng.uniform(0.01, 3.1, 298), [3.1]]))
    ks = np.sort(rng.uniform(1.0, 500.0, 64))
    jl = SphericalBesselCache(ls, xmax=max(20 * ls[-1], ks[-1] * 3.1 * 1.01))
    S = rng.standard_normal((len(taus), len(ks)))
    g = los_integrate(S, ls, taus, ks, jl)
    o = los_oracle(S, taus, ks, jl)
    assert np.max(np.abs(g - o)) <= 1e-12 * np.abs(o).max()
