"""Spherical Bessel function cache for line-of-sight integration.

Restates `SphericalBesselCache` (src/observables/angular.jl:9-48): j_l(x) and j_l'(x) tabulated on
a uniform x grid from 0 to xmax = 20 l_max with spacing close to 2π/15, padded with one duplicate
point, read by cubic Hermite interpolation.  The table is stored [x][l] (l contiguous) as the
reference does, which is also the coalesced order for the GPU kernel (one thread per l).
Values come from scipy.special.spherical_jn (the reference uses Bessels.jl); j_l' uses the
recurrence j_l' = l/(2l+1) j_{l-1} - (l+1)/(2l+1) j_{l+1} (angular.jl:60).
"""
import math

import numpy as np
from scipy.special import spherical_jn


class SphericalBesselCache:
    def __init__(self, ls, xmax=None, dx=2 * math.pi / 15):
        self.l = np.asarray(ls)
        if xmax is None:
            xmax = 20.0 * float(self.l[-1])
        n = int(xmax / dx)
        xs = np.linspace(0.0, xmax, n)
        self.dx = xs[1] - xs[0]
        self.x = np.concatenate([xs, xs[-1:]])
        L = self.l.astype(np.float64)[None, :]
        X = self.x[:, None]
        self.y = np.ascontiguousarray(spherical_jn(self.l[None, :], X))
        jm = np.where(self.l[None, :] > 0, spherical_jn(np.maximum(self.l - 1, 0)[None, :], X), 0.0)
        jp = spherical_jn(self.l[None, :] + 1, X)
        self.dy = np.ascontiguousarray(np.where(self.l[None, :] > 0, L / (2 * L + 1) * jm - (L + 1) / (2 * L + 1) * jp, -jp))

    def __call__(self, il, x):
        """Hermite-interpolated j_l(x) for the il-th tabulated l (angular.jl:40-50)."""
        w = np.asarray(x) / self.dx
        i = np.minimum(w.astype(np.int64), len(self.x) - 2)
        w = w - i
        wm1 = w - 1.0
        return ((1 + 2 * w) * wm1 * wm1 * self.y[i, il] + w * w * (3 - 2 * w) * self.y[i + 1, il] +
                w * wm1 * (wm1 * self.dy[i, il] + w * self.dy[i + 1, il]) * self.dx)
