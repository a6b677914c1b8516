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
