"""The oracle's MLP gradient against central finite differences (CPU)."""
import numpy as np
import pytest


def test_oracle_grad_matches_finite_differences(orc):
    rng = np.random.default_rng(6)
    h1 = h2 = 8
    p = (rng.standard_normal(orc.mlp_nparams(h1, h2)) * 0.5).astype(np.float32)
    x = rng.random((64, 8)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]) + x[:, 1], np.cos(2 * x[:, 2]) - x[:, 3]], 1)
    t = t.astype(np.float32)
    _, g = orc.mlp_loss_grad(p, x, t, None, h1, h2)
    eps = 1e-3
    for i in rng.choice(len(p), 24, replace=False):
        e = np.zeros_like(p)
        e[i] = eps
        lp, _ = orc.mlp_loss_grad(p + e, x, t, None, h1, h2)
        lm, _ = orc.mlp_loss_grad(p - e, x, t, None, h1, h2)
        assert g[i] == pytest.approx((lp - lm) / (2 * eps), rel=3e-2, abs=3e-4)


def test_oracle_adam_step_closed_form(orc):
    # first Adam step moves every parameter by lr * sign(grad) (bias-corrected)
    p = np.zeros(5, np.float32)
    g = np.array([1.0, -2.0, 0.5, -0.25, 3.0], np.float32)
    m = np.zeros(5, np.float32)
    v = np.zeros(5, np.float32)
    orc.mlp_adam(p, g, m, v, 0.1, 1)
    np.testing.assert_allclose(p, -0.1 * np.sign(g), rtol=1e-5)


def numpy_forward(p, x, h1, h2):
    """8 -> h1 -> h2 -> 2 ReLU network written as matrix products from the
    documented parameter layout (W1[h1, 8] b1[h1] W2[h2, h1] b2[h2]
    W3[2, h2] b3[2], row-major), independently of the oracle's loops."""
    o = 0
    W1 = p[o:o + h1 * 8].reshape(h1, 8); o += h1 * 8
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h2 * h1].reshape(h2, h1); o += h2 * h1
    b2 = p[o:o + h2]; o += h2
    W3 = p[o:o + 2 * h2].reshape(2, h2); o += 2 * h2
    b3 = p[o:o + 2]; o += 2
    assert o == len(p)
    a1 = np.maximum(x.astype(np.float64) @ W1.T + b1, 0.0)
    a2 = np.maximum(a1 @ W2.T + b2, 0.0)
    return a2 @ W3.T + b3


@pytest.mark.parametrize("h1,h2", [(32, 32), (16, 8), (8, 24)])
def test_oracle_forward_matches_numpy_matmul(orc, h1, h2):
    # unequal hidden widths: a transposed or misplaced weight block changes
    # the shapes the products need, so it cannot agree by accident
    rng = np.random.default_rng(h1 * 100 + h2)
    p = (rng.standard_normal(orc.mlp_nparams(h1, h2)) * 0.7).astype(np.float32)
    x = (rng.standard_normal((200, 8)) * 2).astype(np.float32)
    np.testing.assert_allclose(orc.mlp_forward(p, x, h1, h2), numpy_forward(p, x, h1, h2),
                               rtol=1e-5, atol=1e-5)


def test_oracle_forward_closed_form(orc):
    # identity hidden layers: y0 = sum of relu(x) + 1, y1 = -(relu(x0) + 2 relu(x7))
    h1 = h2 = 8
    p = np.zeros(orc.mlp_nparams(h1, h2), np.float32)
    o = 0
    p[o:o + 64] = np.eye(8, dtype=np.float32).ravel(); o += 64 + 8
    p[o:o + 64] = np.eye(8, dtype=np.float32).ravel(); o += 64 + 8
    p[o:o + 8] = 1.0
    p[o + 8] = -1.0
    p[o + 15] = -2.0
    o += 16
    p[o] = 1.0
    x = np.array([[1, -2, 3, -4, 5, -6, 7, 8], [-1, 0, 0, 0, 0, 0, 0, -3]], np.float32)
    y = orc.mlp_forward(p, x, h1, h2)
    np.testing.assert_allclose(y, [[1 + 3 + 5 + 7 + 8 + 1, -(1 + 16)], [1, 0]], rtol=1e-6)
