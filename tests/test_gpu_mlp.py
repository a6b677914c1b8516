"""Direction-cost network: device kernels against the CPU oracle.
Tolerances: fp32 kernels 1e-4 relative (summation order differs); the bf16
MFMA inference path 3e-2 absolute on outputs of order 1 (bf16 keeps 8 bits
of mantissa)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_params(orc, seed, h1=32, h2=32, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(orc.mlp_nparams(h1, h2)) * scale).astype(np.float32)


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 8)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]) + x[:, 1], np.cos(2 * x[:, 2]) - x[:, 3]], 1)
    return x, t.astype(np.float32)


@pytest.mark.parametrize("h1,h2", [(32, 32), (16, 8)])
def test_infer_matches_oracle(ctx, orc, h1, h2):
    p = random_params(orc, 1, h1, h2)
    x, _ = data(1000, 2)
    ctx.mlp_set(p, h1, h2)
    ref = orc.mlp_forward(p, x, h1, h2)
    np.testing.assert_allclose(ctx.mlp_infer(x, mfma=False), ref, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(ctx.mlp_infer(x, mfma=True), ref, rtol=3e-2, atol=3e-2)


def test_grad_matches_oracle(ctx, orc):
    p = random_params(orc, 3)
    x, t = data(777, 4)
    mask = (np.random.default_rng(5).random((777, 2)) > 0.3).astype(np.float32)
    ctx.mlp_set(p)
    for mk in (None, mask):
        loss, g = ctx.mlp_loss_grad(x, t, mk)
        rloss, rg = orc.mlp_loss_grad(p, x, t, mk)
        assert loss == pytest.approx(rloss, rel=1e-4)
        np.testing.assert_allclose(g, rg, rtol=1e-3, atol=1e-5 * np.abs(rg).max())


def test_device_training_matches_oracle_adam(ctx, orc):
    p = random_params(orc, 9, scale=0.3)
    x, t = data(512, 10)
    ctx.mlp_set(p)
    loss_dev = ctx.mlp_train(x, t, epochs=20, lr=1e-2)
    q = p.copy()
    m = np.zeros_like(q)
    v = np.zeros_like(q)
    for step in range(1, 21):
        _, g = orc.mlp_loss_grad(q, x, t)
        orc.mlp_adam(q, g, m, v, 1e-2, step)
    loss_ref, _ = orc.mlp_loss_grad(q, x, t)
    got, _, _ = ctx.mlp_get()
    # Adam divides by sqrt(v): a parameter whose gradient is ~0 can move by up
    # to lr per step on either side, so compare the bulk of the parameters and
    # the loss rather than every element.
    assert np.mean(np.abs(got - q) <= 1e-3 + 1e-3 * np.abs(q)) > 0.97
    assert loss_dev == pytest.approx(loss_ref, rel=2e-2)
    first, _ = orc.mlp_loss_grad(p, x, t)
    assert loss_dev < first
