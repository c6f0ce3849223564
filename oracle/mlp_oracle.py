"""CPU restatement of the surrogate MLP's forward and backward pass (numpy, FP64 arithmetic with
BF16 rounding at the same points as the GPU kernels) — TEST INFRASTRUCTURE ONLY: imported by
tests/ as the checker of the GPU gradients.  No reference counterpart: the surrogate is part of the
build's own design.

Forward:  x_0 = bf16(x);  z_l = x_l W_l + b_l;  x_{l+1} = bf16(GELU(z_l));  y = z_L (FP32).
Loss:     mean((y - t)²);  dy = 2 (y - t) / n.
Backward: db_l = Σ_rows bf16(d_l);  dW_l = x_lᵀ bf16(d_l)  (both from the BF16 operand tiles);
          d_{l-1} = bf16((bf16(d_l) W_lᵀ) ⊙ GELU'(z_{l-1})).
GELU(x) = x Φ(x) with the normal CDF (math.erf), GELU'(x) = Φ(x) + x φ(x)."""
import math

import numpy as np

_erf = np.vectorize(math.erf)


EXACT = False   # True: no BF16/FP32 rounding anywhere (for finite-difference checks of the oracle)


def bf16(a):
    """Round FP32 values to BF16 (nearest, ties to even) and return them as float64."""
    if EXACT:
        return np.asarray(a, dtype=np.float64)
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def unpack(params, widths):
    L = len(widths) - 1
    W, b, o = [], [], 0
    for l in range(L):
        n = widths[l] * widths[l + 1]
        W.append(np.asarray(params[o:o + n], dtype=np.float64).reshape(widths[l], widths[l + 1]))
        o += n
    for l in range(L):
        b.append(np.asarray(params[o:o + widths[l + 1]], dtype=np.float64))
        o += widths[l + 1]
    return W, b


def forward(params, widths, x):
    W, b = unpack(params, widths)
    Wh = [bf16(w) for w in W]
    xs, zs = [bf16(x)], []
    for l in range(len(W)):
        z = xs[-1] @ Wh[l] + b[l]
        zs.append(z if EXACT else np.asarray(z, dtype=np.float32).astype(np.float64))
        if l < len(W) - 1:
            xs.append(bf16(gelu(zs[-1])))
    return zs[-1], xs, zs, Wh


def loss_and_grad(params, widths, x, t):
    y, xs, zs, Wh = forward(params, widths, x)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    d = 2.0 * (y - t) / n
    L = len(Wh)
    gW, gb = [None] * L, [None] * L
    dh = bf16(d)
    for l in range(L - 1, -1, -1):
        gb[l] = dh.sum(axis=0)
        gW[l] = xs[l].T @ dh
        if l == 0:
            break
        dh = bf16((dh @ Wh[l].T) * gelu_d(zs[l - 1]))
    return loss, np.concatenate([g.ravel() for g in gW] + [g.ravel() for g in gb])
