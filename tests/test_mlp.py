"""Surrogate MLP: the CPU restatement checked by finite differences (CPU), and the GPU kernels
(MFMA BF16 GEMMs, fused GELU epilogues, backward, Adam) checked against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _problem(widths, B, seed):
    rng = np.random.default_rng(seed)
    n = sum(widths[l] * widths[l + 1] for l in range(len(widths) - 1)) + sum(widths[1:])
    p = rng.normal(0.0, 0.3, n)
    return p, rng.uniform(-1, 1, (B, widths[0])), rng.normal(0.0, 1.0, (B, widths[-1]))


def test_oracle_gradient_by_finite_differences():
    widths = [3, 5, 4, 2]
    p, x, t = _problem(widths, 6, 0)
    mo.EXACT = True
    try:
        _, g = mo.loss_and_grad(p, widths, x, t)
        h = 1e-6
        for i in range(0, len(p), 3):
            e = np.zeros_like(p)
            e[i] = h
            fd = (mo.loss_and_grad(p + e, widths, x, t)[0] - mo.loss_and_grad(p - e, widths, x, t)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), i
    finally:
        mo.EXACT = False


def test_oracle_bf16_rounding():
    # BF16 keeps 7 fraction bits: neighbours of 1 are 1 and 1 + 2^-7
    x = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-9, 1 + 3 * 2**-9, -2.5, 3.0e-20], dtype=np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie between 1 and 1 + 2^-7: to the even one
    assert r[2] == 1 + 2**-6                # tie between 1 + 2^-7 and 1 + 2^-6: to the even one
    assert r[3] == 1.0                      # below the halfway point
    assert r[4] == 1 + 2**-7                # above the halfway point
    assert r[5] == -2.5
    assert abs(r[6] - 3.0e-20) <= 3.0e-20 * 2**-8


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    from symboltz.mlp import Mlp
    widths = [7, 96, 80, 40]
    net = Mlp(widths, max_batch=128, seed=3)
    p = net.get_params()
    _, x, t = _problem(widths, 100, 1)
    y, _ = net.forward(x)
    y_ref = mo.forward(p, widths, x)[0]
    assert np.max(np.abs(y - y_ref)) <= 2e-2 * np.max(np.abs(y_ref))
    g, loss = net.grad(x, t)
    loss_ref, g_ref = mo.loss_and_grad(p, widths, x, t)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    assert np.linalg.norm(g - g_ref) <= 2e-2 * np.linalg.norm(g_ref)
    net.close()


@pytest.mark.gpu
def test_gpu_gradient_at_training_size_matches_oracle():
    # batch 1024 through 512-wide layers: the training products take the 16-byte-load kernel with
    # 128-deep contraction steps, the weight gradients split along the batch and accumulated
    # atomically with the bias gradients from the same tiles; the 7-wide input layer takes the
    # small-contraction and small-output kernels
    from symboltz.mlp import Mlp
    print("case: MLP gradient 1024 x [7, 512, 512, 256]", flush=True)
    widths = [7, 512, 512, 256]
    net = Mlp(widths, max_batch=1024, seed=6)
    p = net.get_params()
    _, x, t = _problem(widths, 1024, 3)
    for _ in range(2):   # the accumulating products start from zero on every call
        g, loss = net.grad(x, t)
        loss_ref, g_ref = mo.loss_and_grad(p, widths, x, t)
        assert abs(loss - loss_ref) <= 1e-3 * loss_ref
        assert np.linalg.norm(g - g_ref) <= 2e-2 * np.linalg.norm(g_ref)
        nb = sum(widths[1:])
        assert np.linalg.norm(g[-nb:] - g_ref[-nb:]) <= 2e-2 * np.linalg.norm(g_ref[-nb:])   # biases
    net.close()


@pytest.mark.gpu
def test_gpu_large_batch_forward_matches_oracle():
    # 16384 rows: the hidden layers take the 128 x 128 MFMA kernel on the transposed weights
    # (sbz_mlp_gemm_kk); the input layer (K = 7) keeps the general kernel
    from symboltz.mlp import Mlp
    print("case: MLP forward 16384 x [7, 512, 512, 256]", flush=True)
    widths = [7, 512, 512, 256]
    net = Mlp(widths, max_batch=16384, seed=4)
    p = net.get_params()
    _, x, _ = _problem(widths, 16384, 2)
    y, _ = net.forward(x)
    y_ref = mo.forward(p, widths, x)[0]
    assert np.max(np.abs(y - y_ref)) <= 2e-2 * np.max(np.abs(y_ref))
    assert np.linalg.norm(y - y_ref) <= 5e-3 * np.linalg.norm(y_ref)
    # after an update the transposed copy follows the weights
    net.set_params(p * 0.5)
    y2, _ = net.forward(x)
    y2_ref = mo.forward(p * 0.5, widths, x)[0]
    assert np.linalg.norm(y2 - y2_ref) <= 5e-3 * np.linalg.norm(y2_ref)
    net.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_map():
    from symboltz.mlp import Mlp
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (4096, 7)).astype(np.float32)
    k = np.linspace(0, 3, 32)
    T = (np.sin(X[:, :1] * 2 + k) * X[:, 1:2] + np.cos(X[:, 2:3] * k)).astype(np.float32)
    net = Mlp([7, 128, 128, 32], max_batch=512, seed=0)
    net.set_data(X, T)
    first = None
    for ep in range(40):
        loss, _ = net.train(rng.permutation(4096), 512, 2e-3)
        first = first if first is not None else float(loss[0])
    assert float(loss.mean()) < 0.05 * first
    y, _ = net.forward(X[:256])
    assert np.sqrt(np.mean((y - T[:256]) ** 2)) < 0.3 * np.sqrt(np.mean(T[:256] ** 2))
    net.close()
