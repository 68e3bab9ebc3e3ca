"""Pins the oracle's MLP: its gradient against central finite differences
(double precision, no bfloat16 rounding) and Adam against the published
update rule."""
import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import learned


def _data(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 12)).astype(np.float32) * 2 + 1
    y = rng.normal(size=n).astype(np.float32)
    c = (rng.random(n) < 0.3).astype(np.float32)
    norm = np.concatenate([np.ones(12), 2 * np.ones(12), [0.0, 1.0]]).astype(np.float32)
    return f, y, c, norm


def test_gradient_matches_finite_differences():
    f, y, c, norm = _data(48, 1)
    w = learned.init_weights(3)
    w[1024:1088] = np.random.default_rng(5).normal(size=64) * 0.1  # nonzero b1
    loss, g = ffi.mlp_loss_grad(w, norm, f, y, c, bf16=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(np.arange(0, 12 * 64), 15), rng.choice(np.arange(1024, 5378), 35)])
    h = 1e-3
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp, _ = ffi.mlp_loss_grad(wp, norm, f, y, c, bf16=False)
        lm, _ = ffi.mlp_loss_grad(wm, norm, f, y, c, bf16=False)
        # the blob is float32: use the actually applied step
        fd = (lp - lm) / (float(wp[i]) - float(wm[i]))
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(g[i])) + 1e-5, (i, fd, g[i])
    assert np.all(g[12 * 64:16 * 64] == 0.0)  # padded inputs carry no gradient


def test_bf16_emulation_is_close_to_exact():
    f, y, c, norm = _data(256, 2)
    w = learned.init_weights(4)
    o_exact = ffi.mlp_forward(w, norm, f, bf16=False)
    o_bf = ffi.mlp_forward(w, norm, f, bf16=True)
    assert np.abs(o_exact - o_bf).max() < 0.05


def test_adam_matches_update_rule():
    rng = np.random.default_rng(3)
    w = rng.normal(size=100).astype(np.float32)
    m = np.zeros(100, np.float32)
    v = np.zeros(100, np.float32)
    w0 = w.copy()
    g = rng.normal(size=100).astype(np.float32)
    ffi.adam_step(w, m, v, g, 1, lr=1e-2)
    # first step: m_hat = g, v_hat = g^2 -> w -= lr * g / (|g| + eps)
    np.testing.assert_allclose(w, w0 - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-7)


def test_calibrate_viol_rate_weights_realizations_by_size():
    """The plain method's violation rate is per subvolume-step: a batch
    weights each realization by steps x subvolumes."""
    from pliscuda_amd import learned
    stats = [{"violations": 10, "steps": 100}, {"violations": 0, "steps": 50}]
    r = learned.calibrate_viol_rate(stats, [1000, 10], margin=1.0)
    assert r == pytest.approx(10 / (100 * 1000 + 50 * 10))
    one = learned.calibrate_viol_rate([{"violations": 10, "steps": 100}], 1000, margin=2.0)
    assert one == pytest.approx(2 * 10 / (100 * 1000))
    # no observed violations: rate 0, the judge's Poisson bound still allows 2 per step
    assert learned.calibrate_viol_rate([{"violations": 0, "steps": 5}], 100) == 0.0
