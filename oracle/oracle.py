"""ctypes wrapper of the CPU oracle (oracle/sbz_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, and there only as the checker.  The product (symboltz.jl_amd) never imports it.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "sbz_oracle.cpp"
LIB = HERE / "build" / "libsbz_oracle.so"
CMB_SRC = HERE / "cmb_oracle.cpp"
CMB_LIB = HERE / "build" / "libcmb_oracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_L = None


def build(force=False):
    """Compile the oracle libraries with g++ (the perturbation stage and the CMB stage)."""
    LIB.parent.mkdir(exist_ok=True)
    for src, out in ((SRC, LIB), (CMB_SRC, CMB_LIB)):
        if force or not out.exists() or out.stat().st_mtime < src.stat().st_mtime:
            cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-shared", "-fPIC", "-pthread", str(src), "-o",
                   str(out)]
            subprocess.run(cmd, check=True)
    return LIB


def lib():
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.ora_n_unknowns.argtypes = [_ip]
        L.ora_layout.argtypes = [_ip, _ip]
        sp = [_dp, _dp, _dp, ctypes.c_int]
        L.ora_rhs.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ora_jacobian.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp]
        L.ora_dfdt.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ora_initial_state.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp]
        L.ora_background.argtypes = [_dp] + sp + [ctypes.c_double, _dp]
        L.ora_delta_m.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp]
        L.ora_delta_m.restype = ctypes.c_double
        L.ora_sources.argtypes = [_ip, _dp] + sp + [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ora_visibility.argtypes = [_dp] + sp + [ctypes.c_double, _dp]
        L.ora_p0.argtypes = [_dp, ctypes.c_double]
        L.ora_p0.restype = ctypes.c_double
        L.ora_solve.argtypes = [_ip, _dp] + sp + [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_long, _dp, ctypes.c_int, _dp, _dp, _ip, _i64p,
                                                  ctypes.c_int]
        L.ora_solve_sources.argtypes = L.ora_solve.argtypes + [_dp]
        L.ora_solve_sparse.argtypes = L.ora_solve.argtypes + [_dp]
        _L = L
    return _L


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def model_desc(M):
    """int[5] model descriptor: lmax_g, lmax_nu (-1 none), nx (0 none), lmax_h, de (0 Λ, 1 w0wa)."""
    return np.array([M.lmax_g, M.lmax_nu if M.has_nu else -1, M.nx if M.has_h else 0, M.lmax_h,
                     1 if M.de == "w0wa" else 0], dtype=np.int32)


class Oracle:
    """The oracle bound to one model, one parameter vector and one background spline table."""

    def __init__(self, M, P, t, u, du):
        self.M = M
        self.md = model_desc(M)
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.u = np.ascontiguousarray(u, dtype=np.float64)
        self.du = np.ascontiguousarray(du, dtype=np.float64)
        self.L = lib()
        self.n = self.L.ora_n_unknowns(_p(self.md, ctypes.c_int))

    def _sp(self):
        return [_p(self.t), _p(self.u), _p(self.du), len(self.t)]

    def layout(self):
        out = np.zeros(9, dtype=np.int32)
        self.L.ora_layout(_p(self.md, ctypes.c_int), _p(out, ctypes.c_int))
        keys = ["Phi", "F", "G", "N", "H", "C", "B", "X", "n"]
        return dict(zip(keys, out.tolist()))

    def rhs(self, k, tau, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.n)
        self.L.ora_rhs(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(y), _p(out))
        return out

    def jacobian(self, k, tau):
        J = np.zeros((self.n, self.n))
        self.L.ora_jacobian(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(J))
        return J

    def dfdt(self, k, tau, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.n)
        self.L.ora_dfdt(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(y), _p(out))
        return out

    def initial_state(self, k, tau):
        out = np.zeros(self.n)
        self.L.ora_initial_state(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(out))
        return out

    def background(self, tau):
        out = np.zeros(15)
        self.L.ora_background(_p(self.P), *self._sp(), tau, _p(out))
        return dict(zip(["a", "H", "kappadot", "cs2", "rho_g", "rho_nu", "rho_c", "rho_b", "rho_h", "P_h", "kappa",
                         "rho_x", "w_x", "Xe", "Tb"], out.tolist()))

    def delta_m(self, k, tau, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        return self.L.ora_delta_m(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(y))

    def sources(self, k, tau, y):
        """(S_T, S_E, S_ψ) at (τ, y)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(3)
        self.L.ora_sources(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), k, tau, _p(y), _p(out))
        return out

    def visibility(self, tau):
        """(v, v', v'', e^{-κ}) at τ, as the sources use them."""
        out = np.zeros(4)
        self.L.ora_visibility(_p(self.P), *self._sp(), tau, _p(out))
        return out

    def p0(self, k):
        return self.L.ora_p0(_p(self.P), k)

    def solve(self, ks, tstart, rtol=1e-5, atol=1e-5, maxiters=100000, saveat=None, nthreads=None, sources=False,
              sparse=False):
        """sparse: Jacobian by column coloring and a static-pattern sparse LU (the CPU baseline's
        solver) instead of the dense J by unit vectors and the dense partial-pivot LU."""
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        tstart = np.ascontiguousarray(np.broadcast_to(tstart, ks.shape), dtype=np.float64)
        nm = len(ks)
        if saveat is None:
            saveat = np.zeros(0)
        saveat = np.ascontiguousarray(saveat, dtype=np.float64)
        ns = len(saveat)
        usave = np.zeros((nm, ns, self.n)) if ns else np.zeros(1)
        ufinal = np.zeros((nm, self.n))
        ret = np.zeros(nm, dtype=np.int32)
        stats = np.zeros((nm, 6), dtype=np.int64)
        if nthreads is None:
            nthreads = os.cpu_count() or 1
        src = np.zeros((nm, ns, 3)) if (sources and ns) else None
        fn = self.L.ora_solve_sparse if sparse else self.L.ora_solve_sources
        fn(_p(self.md, ctypes.c_int), _p(self.P), *self._sp(), _p(ks), _p(tstart), nm, rtol,
                                 atol, maxiters, _p(saveat), ns, _p(usave), _p(ufinal), _p(ret, ctypes.c_int),
                                 _p(stats, ctypes.c_int64), nthreads, _p(src) if src is not None else None)
        return dict(u=ufinal, usave=usave if ns else None, retcode=ret, stats=stats, src=src)

    def spectrum_matter(self, ks, kt_ini=1e-2, tau_ini_max=1e-4, **kw):
        """P(k) at τ0 as spectrum_matter(prob, ks) (fourier.jl:79-83,90-97)."""
        ks = np.asarray(ks, dtype=np.float64)
        t0, t1 = self.t[0], self.t[-1]
        tstart = np.clip(np.minimum(kt_ini / ks, tau_ini_max), t0, t1)
        r = self.solve(ks, tstart, **kw)
        dm = np.array([self.delta_m(k, t1, r["u"][i]) for i, k in enumerate(ks)])
        P = np.array([self.p0(k) for k in ks]) * dm**2
        return P, r
