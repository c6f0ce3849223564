"""CPU oracle of the line-of-sight integral (TEST INFRASTRUCTURE ONLY).

Restates los_integrate (src/observables/angular.jl:109-185, non-Limber branch) in numpy:
trapezoid weights (:124-129), χ = τ0 - τ, I_l(k) = Σ_τ w_τ S(τ,k) j_l(kχ) with the Bessel cache's
Hermite interpolation (:40-50)."""
import numpy as np


def los_integrate(Ss, taus, ks, jl):
    taus = np.asarray(taus, dtype=np.float64)
    n = len(taus)
    w = np.empty(n)
    w[0] = 0.5 * (taus[1] - taus[0])
    w[1:-1] = 0.5 * (taus[2:] - taus[:-2])
    w[-1] = 0.5 * (taus[-1] - taus[-2])
    chi = taus[-1] - taus
    out = np.zeros((len(ks), len(jl.l)))
    for ik, k in enumerate(ks):
        x = k * chi
        for il in range(len(jl.l)):
            out[ik, il] = np.sum(w * Ss[:, ik] * jl(il, x))
    return out
