"""Pin the Rodas5P tableau of csrc/sbz_rodas5p.h (shared by the background solver and the GPU
kernel) and the oracle's copy: no copy of the published coefficients exists offline, so they are
verified by what the method must satisfy — local order 6 (global 5) of the solution, order 5 of
the embedded solution, consistency of c_i and d_i, and L-stability / stiff accuracy."""
import re
from pathlib import Path

import mpmath as mp
import pytest

ROOT = Path(__file__).resolve().parent.parent
HDR = ROOT / "symboltz.jl_amd" / "csrc" / "sbz_rodas5p.h"


def load_tableau():
    vals = {}
    for m in re.finditer(r"#define SBZ_R5_(\w+)\s+\(?([-0-9.eE]+)\)?", HDR.read_text()):
        vals[m.group(1)] = mp.mpf(m.group(2))
    a = {(i, j): vals[f"A{i}{j}"] for i in range(2, 7) for j in range(1, i) if f"A{i}{j}" in vals}
    for j in range(1, 6):
        a[(7, j)] = a[(6, j)]
        a[(8, j)] = a[(6, j)]
    a[(7, 6)] = a[(8, 6)] = a[(8, 7)] = mp.mpf(1)
    C = {(i, j): vals[f"C{i}{j}"] for i in range(2, 9) for j in range(1, i)}
    c = [0, 0, vals["C2"], vals["C3"], vals["C4"], vals["C5"], 1, 1, 1]
    d = [0, vals["D1"], vals["D2"], vals["D3"], vals["D4"], vals["D5"], 0, 0, 0]
    return vals["GAMMA"], a, C, c, d


def step(tab, f, jac, dfdt, t, y, h):
    g, a, C, c, d = tab
    n = len(y)
    W = mp.eye(n) / (g * h) - jac(t, y)
    ft = dfdt(t, y)
    k = [None] * 9
    for i in range(1, 9):
        u = y.copy()
        for j in range(1, i):
            u += a[(i, j)] * k[j]
        rhs = f(t + c[i] * h, u) + d[i] * h * ft
        for j in range(1, i):
            rhs += (C[(i, j)] / h) * k[j]
        k[i] = mp.lu_solve(W, rhs)
    return u + k[8], u


def test_convergence_order():
    mp.mp.dps = 40
    tab = load_tableau()
    f = lambda t, y: mp.matrix([-y[1] + t * y[0] * y[1], y[0] - mp.mpf("0.3") * y[1] ** 2 + mp.sin(t)])
    jac = lambda t, y: mp.matrix([[t * y[1], -1 + t * y[0]], [1, -mp.mpf("0.6") * y[1]]])
    dfdt = lambda t, y: mp.matrix([y[0] * y[1], mp.cos(t)])
    t0, y0 = mp.mpf("0.3"), mp.matrix([mp.mpf("0.5"), mp.mpf("-0.2")])
    ex = mp.odefun(lambda t, y: [-y[1] + t * y[0] * y[1], y[0] - mp.mpf("0.3") * y[1] ** 2 + mp.sin(t)], t0,
                   [y0[0], y0[1]])
    errs, eerrs = [], []
    for p in range(3, 7):
        h = mp.mpf(2) ** (-p)
        y1, yh = step(tab, f, jac, dfdt, t0, y0, h)
        ye = mp.matrix(ex(t0 + h))
        errs.append(mp.norm(y1 - ye))
        eerrs.append(mp.norm(yh - ye))
    orders = [float(mp.log(errs[i] / errs[i + 1], 2)) for i in range(len(errs) - 1)]
    eorders = [float(mp.log(eerrs[i] / eerrs[i + 1], 2)) for i in range(len(eerrs) - 1)]
    assert all(o > 5.7 for o in orders), orders          # local error O(h^6): global order 5
    assert all(4.7 < o < 5.5 for o in eorders), eorders  # embedded: local O(h^5)


def test_consistency_and_stiff_accuracy():
    g, a, C, c, d = load_tableau()
    # c_i = Σ_j α_ij and d_i = Σ_j γ_ij - γ after transforming back from (a, C) to (α, Γ)
    Ginv = mp.matrix(8, 8)
    for i in range(8):
        Ginv[i, i] = 1 / g
        for j in range(i):
            Ginv[i, j] = -C.get((i + 1, j + 1), 0)
    Gam = mp.inverse(Ginv)
    A = mp.matrix(8, 8)
    for i in range(8):
        for j in range(i):
            A[i, j] = a.get((i + 1, j + 1), 0)
    alpha = A * Gam
    for i in range(8):
        assert abs(sum(alpha[i, j] for j in range(8)) - c[i + 1]) < 1e-12
        gi = sum(Gam[i, j] for j in range(8))
        if i < 5:
            assert abs(gi - d[i + 1]) < 1e-12


@pytest.mark.parametrize("lam", [-1e6, -1e10])
def test_L_stability(lam):
    mp.mp.dps = 30
    tab = load_tableau()
    f = lambda t, y: mp.matrix([lam * y[0]])
    jac = lambda t, y: mp.matrix([[lam]])
    dfdt = lambda t, y: mp.matrix([0])
    y1, _ = step(tab, f, jac, dfdt, mp.mpf(0), mp.matrix([mp.mpf(1)]), mp.mpf(1))
    assert abs(y1[0]) < 1e3 / abs(lam)   # R(z) → 0 as z → -∞
