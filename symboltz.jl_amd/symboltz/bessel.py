"""Spherical Bessel function cache for line-of-sight integration.

Restates `SphericalBesselCache` (src/observables/angular.jl:9-48): j_l(x) and j_l'(x) tabulated on
a uniform x grid from 0 to xmax = 20 l_max with spacing close to 2π/15, padded with one duplicate
point, read by cubic Hermite interpolation.  The table is stored [x][l] (l contiguous) as the
reference does, which is also the coalesced order for the GPU kernels (threads walk l).

The values are computed on the GPU (`sbz_bessel_table`, three-term recurrence with Miller's
downward form where l exceeds x) the first time `y` or `dy` is read; the C_l pipeline
(`symboltz.cmb.spectrum_cmb`) builds the same table on the device itself and never copies it.
j_l' = l/(2l+1) j_{l-1} - (l+1)/(2l+1) j_{l+1} as at angular.jl:60.
"""
import ctypes
import math

import numpy as np

from . import _lib


class SphericalBesselCache:
    def __init__(self, ls, xmax=None, dx=2 * math.pi / 15):
        self.l = np.asarray(ls)
        if self.l.ndim != 1 or len(self.l) == 0 or np.any(np.diff(self.l) <= 0) or self.l[0] < 0:
            raise ValueError("ls must be a non-empty, strictly ascending vector of non-negative integers")
        if xmax is None:
            xmax = 20.0 * float(self.l[-1])
        n = int(xmax / dx)                       # trunc(Int, (xmax - xmin) / dx) points from 0 to xmax
        self.dx = xmax / (n - 1)                 # the resulting step, as the reference keeps it
        self.n_x = n + 1                         # plus the duplicate padding point
        self.xmax = xmax
        self._y = self._dy = None
        self._ctx = None
        self.context_ms = 0.0

    def context(self):
        """Device C_l context (`sbz_cmb_create`): the table built once on the GPU and kept there."""
        if self._ctx is None:
            import time
            L = _lib.lib()
            L.sbz_cmb_create.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_void_p)]
            L.sbz_cmb_destroy.argtypes = [ctypes.c_void_p]
            ls = np.ascontiguousarray(self.l, dtype=np.int32)
            h = ctypes.c_void_p()
            t = time.perf_counter()
            _lib.check(L.sbz_cmb_create(_lib.ptr(ls, ctypes.c_int), len(ls), self.n_x, self.dx, ctypes.byref(h)),
                       "sbz_cmb_create")
            self.context_ms = (time.perf_counter() - t) * 1e3
            self._ctx = h
        return self._ctx

    def close(self):
        if getattr(self, "_ctx", None) is not None:
            _lib.lib().sbz_cmb_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def x(self):
        x = np.arange(self.n_x, dtype=np.float64) * self.dx
        x[-1] = x[-2]
        return x

    def _table(self):
        if self._y is None:
            L = _lib.lib()
            dp = ctypes.POINTER(ctypes.c_double)
            L.sbz_bessel_table.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           dp, dp]
            ls = np.ascontiguousarray(self.l, dtype=np.int32)
            y = np.zeros((self.n_x, len(ls)))
            dy = np.zeros((self.n_x, len(ls)))
            _lib.check(L.sbz_bessel_table(_lib.ptr(ls, ctypes.c_int), len(ls), self.n_x, self.dx, _lib.ptr(y),
                                          _lib.ptr(dy)), "sbz_bessel_table")
            self._y, self._dy = y, dy
        return self._y, self._dy

    @property
    def y(self):
        return self._table()[0]

    @property
    def dy(self):
        return self._table()[1]

    def __call__(self, il, x):
        """Hermite-interpolated j_l(x) for the il-th tabulated l (angular.jl:40-50)."""
        y, dy = self._table()
        w = np.asarray(x) / self.dx
        i = np.minimum(w.astype(np.int64), self.n_x - 2)
        w = w - i
        wm1 = w - 1.0
        return ((1 + 2 * w) * wm1 * wm1 * y[i, il] + w * w * (3 - 2 * w) * y[i + 1, il] +
                w * wm1 * (wm1 * dy[i, il] + w * dy[i + 1, il]) * self.dx)

    def __repr__(self):
        return (f"jₗ(x) Hermite interpolation cache for {self.l[0]} ≤ l ≤ {self.l[-1]} and "
                f"0.0 ≤ x ≤ {self.x[-1]} ({self.n_x * len(self.l) * 16 / 2**20:.1f} MiB)")
