"""Momentum quadrature for massive neutrinos.

Restates `momentum_quadrature` (src/models/neutrinos.jl:55-60): an N-point Gaussian rule for
∫_0^∞ dx x² f0(x) g(x), f0 = 1/(eˣ+1), built in the substituted variable u = 1/(1 + x/100)
on (0, 1) with weight w(u) = |dx/du| x(u)² f0(x(u)), exactly as QuadGK's `gauss(W, N, a, b)`
does (Gauss rule exact for polynomials of degree 2N-1 in u).  The nodes are returned as
x_i = 100 (1 - u_i)/u_i together with the weights W_i.

The recurrence coefficients are obtained with the discretised Stieltjes procedure on a
high-order piecewise Gauss–Legendre discretisation of w(u) du, then Golub–Welsch.
"""
import numpy as np

L_SCALE = 100.0


def _x_of_u(u):
    return L_SCALE * (1.0 - u) / u


def _weight(u):
    x = _x_of_u(u)
    dxdu = L_SCALE / u**2
    # x² f0(x) with f0 = 1/(e^x + 1), written to avoid overflow for large x
    return dxdu * x * x * np.exp(-x) / (1.0 + np.exp(-x))


def _discretisation(npts=200):
    # Break points concentrated where the weight lives (x between 0 and ~60)
    xb = np.array([0.0, 0.5, 1, 2, 3, 4, 6, 8, 11, 15, 20, 27, 35, 45, 60, 80, 110, 160, 250, 500, 2000, 1e5])
    ub = np.sort(1.0 / (1.0 + xb / L_SCALE))
    ub = np.concatenate([[0.0], ub])
    gx, gw = np.polynomial.legendre.leggauss(npts)
    us, ws = [], []
    for lo, hi in zip(ub[:-1], ub[1:]):
        u = 0.5 * (hi - lo) * gx + 0.5 * (hi + lo)
        w = 0.5 * (hi - lo) * gw
        us.append(u)
        ws.append(w * _weight(u))
    return np.concatenate(us), np.concatenate(ws)


def momentum_quadrature(n):
    """Return (x, W): n nodes and weights with Σ W_i g(x_i) ≈ ∫_0^∞ x² g(x)/(eˣ+1) dx."""
    if n < 1:
        raise ValueError("need at least one momentum bin")
    u, w = _discretisation()
    # Stieltjes procedure for the monic recurrence p_{k+1} = (u - α_k) p_k - β_k p_{k-1}
    alpha = np.zeros(n)
    beta = np.zeros(n)
    p_prev = np.zeros_like(u)
    p = np.ones_like(u)
    norm_prev = 1.0
    for k in range(n):
        nrm = np.sum(w * p * p)
        alpha[k] = np.sum(w * u * p * p) / nrm
        beta[k] = nrm if k == 0 else nrm / norm_prev
        p_next = (u - alpha[k]) * p - (beta[k] * p_prev if k > 0 else 0.0)
        p_prev, p, norm_prev = p, p_next, nrm
    # Golub–Welsch
    J = np.diag(alpha) + np.diag(np.sqrt(beta[1:]), 1) + np.diag(np.sqrt(beta[1:]), -1)
    nodes, vecs = np.linalg.eigh(J)
    weights = beta[0] * vecs[0, :] ** 2
    x = _x_of_u(nodes)
    # order by increasing u, as QuadGK returns them (i.e. decreasing x)
    order = np.argsort(nodes)
    return x[order], weights[order]
