import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "symboltz.jl_amd"))
sys.path.insert(0, str(ROOT))

from symboltz import _lib, models  # noqa: E402
from oracle import oracle  # noqa: E402

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


class Setup:
    """A model, its Planck18 parameters, the product background solve and the oracle on it."""

    def __init__(self, M, pars=None):
        self.M = M
        self.pars = pars if pars is not None else models.parameters_Planck18(M)
        P = models.make_params(M, self.pars)
        self.P, self.t, self.u, self.du, _ = _lib.background_solve(P)
        self.O = oracle.Oracle(M, self.P, self.t, self.u, self.du)
        self.knots = np.concatenate([self.t, self.u.ravel(), self.du.ravel()])


@pytest.fixture(scope="session")
def fixture5():
    """The reference test-suite's model: ΛCDM(lmax = 5) with Planck18 (test/runtests.jl:14-16)."""
    return Setup(models.LCDM(lmax=5))


@pytest.fixture(scope="session")
def massless10():
    return Setup(models.LCDM(lmax=10, h=False))


class Probe:
    """ctypes view of the host build of a model's emitted device functions."""

    def __init__(self, M):
        from symboltz import build
        path = build.build_probe(M)
        L = ctypes.CDLL(str(path))
        dp = ctypes.POINTER(ctypes.c_double)
        L.sbz_probe_bg.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double, dp, dp]
        L.sbz_probe_f.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.sbz_probe_w.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.sbz_probe_lu.argtypes = [dp]
        L.sbz_probe_solve.argtypes = [dp, dp]
        L.sbz_probe_dfdt.argtypes = [dp, dp, dp, ctypes.c_double, dp, dp]
        L.sbz_probe_ic.argtypes = [dp, dp, ctypes.c_double, dp]
        L.sbz_probe_dm.argtypes = [dp, dp, ctypes.c_double, dp]
        L.sbz_probe_dm.restype = ctypes.c_double
        L.sbz_probe_src.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        self.L = L
        self.n = L.sbz_probe_n()
        self.nb = L.sbz_probe_nb()
        self.nnz = L.sbz_probe_nnz()

    @staticmethod
    def p(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def bg(self, S, tau):
        B, dB = np.zeros(self.nb), np.zeros(self.nb)
        self.L.sbz_probe_bg(self.p(S.P), self.p(S.knots), len(S.t), tau, self.p(B), self.p(dB))
        return B, dB

    def f(self, S, k, tau, u):
        B, _ = self.bg(S, tau)
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.zeros(self.n)
        self.L.sbz_probe_f(self.p(B), self.p(S.P), k, self.p(u), self.p(du))
        return du

    def w(self, S, k, tau, ig):
        B, _ = self.bg(S, tau)
        W = np.zeros(self.nnz)
        self.L.sbz_probe_w(self.p(B), self.p(S.P), k, ig, self.p(W))
        return W

    def lu_solve(self, W, b):
        W = W.copy()
        ok = self.L.sbz_probe_lu(self.p(W))
        x = np.ascontiguousarray(b, dtype=np.float64).copy()
        self.L.sbz_probe_solve(self.p(W), self.p(x))
        return ok, x

    def dfdt(self, S, k, tau, u):
        B, dB = self.bg(S, tau)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(self.n)
        self.L.sbz_probe_dfdt(self.p(B), self.p(dB), self.p(S.P), k, self.p(u), self.p(out))
        return out

    def ic(self, S, k, tau):
        B, _ = self.bg(S, tau)
        u = np.zeros(self.n)
        self.L.sbz_probe_ic(self.p(B), self.p(S.P), k, self.p(u))
        return u

    def src(self, S, k, tau, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(3)
        work = np.zeros(4 * self.n)
        self.L.sbz_probe_src(self.p(S.P), self.p(S.knots), len(S.t), k, tau, self.p(u), self.p(out), self.p(work))
        return out

    def dm(self, S, k, tau, u):
        B, _ = self.bg(S, tau)
        u = np.ascontiguousarray(u, dtype=np.float64)
        return self.L.sbz_probe_dm(self.p(B), self.p(S.P), k, self.p(u))


@pytest.fixture(scope="session")
def probe5():
    return Probe(models.LCDM(lmax=5))


@pytest.fixture(scope="session")
def probe10():
    return Probe(models.LCDM(lmax=10, h=False))
