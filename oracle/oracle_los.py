"""CPU oracle of the line-of-sight integral (TEST INFRASTRUCTURE ONLY).

Restates los_integrate (src/observables/angular.jl:109-185, non-Limber branch) in numpy:
trapezoid weights (:124-129), χ = τ0 - τ, I_l(k) = Σ_τ w_τ S(τ,k) j_l(kχ) with the Bessel cache's
Hermite interpolation (:40-50).  The table is the oracle's own (`oracle_cmb.bessel_table`)."""
import numpy as np


def hermite(table, il, x):
    dx, y, dy = table
    w = np.asarray(x) / dx
    i = np.minimum(w.astype(np.int64), y.shape[0] - 2)
    w = w - i
    wm1 = w - 1.0
    return ((1 + 2 * w) * wm1 * wm1 * y[i, il] + w * w * (3 - 2 * w) * y[i + 1, il] +
            w * wm1 * (wm1 * dy[i, il] + w * dy[i + 1, il]) * dx)


def los_integrate(Ss, taus, ks, table):
    taus = np.asarray(taus, dtype=np.float64)
    n = len(taus)
    w = np.empty(n)
    w[0] = 0.5 * (taus[1] - taus[0])
    w[1:-1] = 0.5 * (taus[2:] - taus[:-2])
    w[-1] = 0.5 * (taus[-1] - taus[-2])
    chi = taus[-1] - taus
    nl = table[1].shape[1]
    out = np.zeros((len(ks), nl))
    for ik, k in enumerate(ks):
        x = k * chi
        for il in range(nl):
            out[ik, il] = np.sum(w * Ss[:, ik] * hermite(table, il, x))
    return out
