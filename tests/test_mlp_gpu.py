"""Learned tau on the GPU: MFMA inference and gradients against the oracle
(bfloat16 operands emulated, so the only differences are fp32 summation
order), training on the GPU, records against the oracle's features and Cao
tau', and the learned stepping mode with its judge."""
import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import capi, learned
from pliscuda_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _ctx_small():
    net, topo, x0 = H.small_bruss()
    return net, topo, x0, capi.Context(net, [topo], [x0], seed=3)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    f = (rng.normal(size=(n, 12)) * 2 + 1).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    c = (rng.random(n) < 0.3).astype(np.float32)
    norm = np.concatenate([np.ones(12), 2 * np.ones(12), [0.0, 1.0]]).astype(np.float32)
    return f, y, c, norm


def test_mfma_inference_matches_oracle():
    _, _, _, ctx = _ctx_small()
    f, y, c, norm = _data(1000, 1)
    w = learned.init_weights(2)
    w[1024:1088] = 0.05
    ctx.mlp_set(w, norm)
    out = ctx.mlp_forward(f)
    ref = ffi.mlp_forward(w, norm, f, bf16=True)
    assert np.abs(out - ref).max() < 2e-2, np.abs(out - ref).max()
    assert np.abs(out - ref).mean() < 2e-3
    ctx.close()


def test_mfma_gradient_matches_oracle():
    _, _, _, ctx = _ctx_small()
    f, y, c, norm = _data(512, 2)
    w = learned.init_weights(5)
    ctx.mlp_set(w, norm)
    loss, g = ctx.mlp_loss_grad(f, y, c)
    lref, gref = ffi.mlp_loss_grad(w, norm, f, y, c, bf16=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 3e-2, rel
    # per-layer agreement of the direction
    for a, b in ((0, 1024), (1088, 5184), (5248, 5376)):
        cos = np.dot(g[a:b], gref[a:b]) / (np.linalg.norm(g[a:b]) * np.linalg.norm(gref[a:b]))
        assert cos > 0.995, (a, b, cos)
    ctx.close()


def test_gpu_training_reduces_loss():
    _, _, _, ctx = _ctx_small()
    rng = np.random.default_rng(9)
    f = (rng.normal(size=(20000, 12))).astype(np.float32)
    y = (np.sin(f[:, 0]) + 0.5 * f[:, 5] - 0.3 * f[:, 11]).astype(np.float32)
    c = (f[:, 2] > 0.5).astype(np.float32)
    norm = np.concatenate([np.zeros(12), np.ones(12), [0.0, 1.0]]).astype(np.float32)
    ctx.mlp_set(learned.init_weights(1), norm)
    losses = ctx.mlp_train(f, y, c, steps=400, batch=4096, lr=3e-3, seed=2)
    assert losses[-20:].mean() < 0.3 * losses[:5].mean(), (losses[:5], losses[-5:])
    ctx.close()


def test_records_match_oracle():
    net, topo, x0, ctx = _ctx_small()
    feats, ylog, crit = ctx.records(0)
    m = ffi.Model(net, topo)
    of = ffi.features(m, x0, 0.03, 10, 0.0)
    np.testing.assert_allclose(feats, of, rtol=2e-6, atol=2e-6)  # hardware log on the device
    tp, a0c, _ = ffi.tau_all(m, x0, 0.03, 10, 3)
    L = ffi.lib()
    ref = np.array([L.orc_log(float(t)) if t < 1e6 else L.orc_log(1e6) for t in tp], dtype=np.float32)
    assert np.array_equal(ylog, ref)
    assert np.array_equal(crit, (a0c > 0).astype(np.float32))
    ctx.close()


def test_learned_mode_runs_and_is_judged():
    """Train on records of a plain run, then step with the learned proposal:
    every step is either a kept proposal or a Cao fallback, populations stay
    valid and the trajectory stays close to the plain one."""
    net = W.brusselator_network(W.SEED_BRUSS)
    topo = W.random_mesh(16, 16, 8, 5)
    x0 = W.poisson_state(5, topo["N"], [60, 40, 80, 100, 20], 5)
    rec_ctx = capi.Context(net, [topo], [x0], seed=11)
    F, Y, Cc = learned.collect_records(rec_ctx, snapshots=20, stride=5, per_snapshot=2048)
    plain_stats = [rec_ctx.stats(0)]
    norm, losses = learned.train(rec_ctx, F, Y, Cc, steps=600, batch=8192, lr=2e-3)
    assert losses[-10:].mean() < losses[:3].mean()
    w = rec_ctx.mlp_get()
    vmax = learned.calibrate_viol_rate(plain_stats, topo["N"])
    rec_ctx.close()

    plain = capi.Context(net, [topo], [x0], seed=21)
    plain.run(np.inf, step_limit=150)
    sp = plain.stats()
    lrn = capi.Context(net, [topo], [x0], seed=21)
    lrn.mlp_set(w, norm)
    lrn.set_tau_mode(True, kappa=1.0, theta=3.0, viol_rate_max=vmax)
    lrn.run(sp["t"], step_limit=10**6)
    sl = lrn.stats()
    assert sl["status"] == 1 and sl["t"] == sp["t"]
    assert sl["learned_steps"] > 0
    props = sl["learned_steps"] + sl["learned_rejects"]
    print("learned", sl, "plain", sp)
    assert sl["learned_rejects"] <= 0.25 * props
    xl = lrn.get_state().reshape(5, -1).astype(float).sum(1)
    xp = plain.get_state().reshape(5, -1).astype(float).sum(1)
    assert np.all(np.abs(xl - xp) <= 0.05 * xp + 50), (xl, xp)
    plain.close()
    lrn.close()
