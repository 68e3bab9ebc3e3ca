"""Learned tau on the GPU: MFMA inference and gradients against the oracle
for both activations (bfloat16 operands emulated, so the only differences
are fp32 summation order), training on the GPU, records against the
oracle's features, Cao tau' and trial-leap labels, the proposals fused into
the leap pass against the standalone pass, and the learned stepping mode:
judged, absorbing states, ensemble statistics against the plain method, and
the learned-tau configuration at full size (the 1,048,576-subvolume
Brusselator mesh and the 128-realization uneven batch)."""
import math

import numpy as np
import pytest
from scipy import stats

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


@pytest.mark.parametrize("act", [0, 1])
def test_mfma_inference_matches_oracle(act):
    _, _, _, ctx = _ctx_small()
    f, y, c, norm = _data(1000, 1)
    w = learned.init_weights(2)
    w[1024:1088] = 0.05
    ctx.mlp_set(w, norm)
    ctx.mlp_activation(act)
    out = ctx.mlp_forward(f)
    ref = ffi.mlp_forward(w, norm, f, bf16=True, act=act)
    assert np.abs(out - ref).max() < 2e-2, np.abs(out - ref).max()
    assert np.abs(out - ref).mean() < 2e-3
    ctx.close()


@pytest.mark.parametrize("act", [0, 1])
def test_mfma_gradient_matches_oracle(act):
    _, _, _, ctx = _ctx_small()
    f, y, c, norm = _data(512, 2)
    w = learned.init_weights(5)
    ctx.mlp_set(w, norm)
    ctx.mlp_activation(act)
    loss, g = ctx.mlp_loss_grad(f, y, c)
    lref, gref = ffi.mlp_loss_grad(w, norm, f, y, c, bf16=True, act=act)
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
    """Features to the hardware log's tolerance, log Cao tau' and the trial
    leaps' accept/reject labels bit for bit, at the initial state and after
    seven steps (step index and previous tau in the streams and features)."""
    net, topo, x0, ctx = _ctx_small()
    ctx.set_tau_mode(False, theta=1.0)
    m = ffi.Model(net, topo)
    for steps in (0, 7):
        if steps:
            ctx.run(np.inf, step_limit=steps)
        x = ctx.get_state()
        st = ctx.stats()
        prev = np.float32(np.log(np.float32(st["tau"]))) if st["tau"] > 0 else np.float32(0.0)
        feats, ylog, lab = ctx.records(0)
        of, oy, ol = ffi.records(m, x, 0.03, 10, 3, theta=1.0, step=st["steps"], prev_logtau=float(prev))
        np.testing.assert_allclose(feats, of, rtol=2e-6, atol=2e-6)  # hardware log on the device
        assert np.array_equal(ylog, oy)
        assert np.array_equal(lab, ol), (steps, np.flatnonzero(lab != ol)[:10])
        assert 0.0 < lab.mean() < 1.0
    ctx.close()


def _trained(net, topo, x0, seed=11, snapshots=12, stride=4, per_snapshot=2048, steps=400, batch=8192):
    """Records from a plain run, a short GPU training; returns (w, norm, vmax)."""
    topos = topo if isinstance(topo, list) else [topo]
    xs = x0 if isinstance(topo, list) else [x0]
    rec = capi.Context(net, topos, xs, seed=seed)
    F, Y, L = learned.collect_records(rec, snapshots=snapshots, stride=stride, per_snapshot=per_snapshot)
    plain_stats = [rec.stats(r) for r in range(rec.n)]
    norm, losses = learned.train(rec, F, Y, L, steps=steps, batch=batch, lr=2e-3)
    assert losses[-10:].mean() < losses[:3].mean()
    w = rec.mlp_get()
    vmax = learned.calibrate_viol_rate(plain_stats, [int(t["N"]) for t in topos])
    rec.close()
    return w, norm, vmax


def test_fused_proposals_equal_standalone_pass():
    """The proposals the leap pass makes for the state it produces equal the
    standalone learned tau pass on that state: a run that forces the
    standalone pass every step (setting the state again) is bit-identical."""
    net, topo, x0 = H.small_bruss()
    w, norm, vmax = _trained(net, topo, x0)
    runs = []
    for force in (False, True):
        ctx = capi.Context(net, [topo], [x0], seed=5)
        ctx.mlp_set(w, norm)
        ctx.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=vmax)
        for k in range(1, 31):
            ctx.run(np.inf, step_limit=k)
            if force:
                ctx.set_state(0, ctx.get_state())
        runs.append((ctx.get_state(), ctx.stats()))
        ctx.close()
    (xa, sa), (xb, sb) = runs
    assert np.array_equal(xa, xb)
    for f in ("steps", "learned_steps", "learned_rejects", "leapbacks", "firings", "t", "violations"):
        assert sa[f] == sb[f], (f, sa[f], sb[f])
    assert sa["learned_steps"] > 0


def test_learned_absorbing_state_stops_like_plain():
    """A -> 0 with diffusion: once every molecule is gone no event is
    possible, and the learned mode stops with status 2 like the plain one
    (an infinite proposal where a subvolume has no non-critical channel)."""
    net = W.network(["A"], [({"A": 1}, {})], [2.0], [0.5])
    topo = W.cubic_grid(4, 4, 4, periodic=False)
    x0 = np.full(topo["N"], 30, np.uint32)
    out = []
    for learned_on in (False, True):
        ctx = capi.Context(net, [topo], [x0], seed=8)
        if learned_on:
            norm = np.concatenate([np.zeros(12), np.ones(12), [-6.0, 1.0]]).astype(np.float32)
            ctx.mlp_set(learned.init_weights(3), norm)
            ctx.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=1.0)
        ctx.run(1e6)
        s = ctx.stats()
        assert s["status"] == 2 and s["t"] < 1e6, s
        assert not ctx.get_state().any()
        out.append(s)
        ctx.close()
    assert out[1]["learned_steps"] > 0


def test_learned_mode_runs_and_is_judged():
    """Train on records of a plain run, then step with the learned proposal:
    every step is either a kept proposal or a Cao fallback, populations stay
    valid, few proposals are rejected and the trajectory stays close to the
    plain one."""
    net = W.brusselator_network(W.SEED_BRUSS)
    topo = W.random_mesh(16, 16, 8, 5)
    x0 = W.poisson_state(5, topo["N"], [60, 40, 80, 100, 20], 5)
    w, norm, vmax = _trained(net, topo, x0, snapshots=20, stride=5, steps=600)
    plain = capi.Context(net, [topo], [x0], seed=21)
    plain.run(np.inf, step_limit=150)
    sp = plain.stats()
    lrn = capi.Context(net, [topo], [x0], seed=21)
    lrn.mlp_set(w, norm)
    lrn.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=vmax)
    lrn.run(sp["t"], step_limit=10**6)
    sl = lrn.stats()
    assert sl["status"] == 1 and sl["t"] == sp["t"]
    props = sl["learned_steps"] + sl["learned_rejects"]
    print("learned", sl, "plain", sp)
    assert sl["learned_steps"] > 0.8 * sl["steps"]
    assert sl["learned_rejects"] <= 0.05 * props
    xl = lrn.get_state().reshape(5, -1).astype(float).sum(1)
    xp = plain.get_state().reshape(5, -1).astype(float).sum(1)
    assert np.all(np.abs(xl - xp) <= 0.05 * xp + 50), (xl, xp)
    plain.close()
    lrn.close()


def test_learned_ensemble_matches_plain():
    """64 realizations of a 6x6x4 Brusselator mesh to the same time, plain
    Cao tau and learned tau from independent streams: per-species totals
    agree within 3 standard errors (means and variances) and by two-sample
    KS (p > 0.01); proposals are rarely rejected."""
    net = W.brusselator_network(W.SEED_BRUSS)
    topo = W.random_mesh(6, 6, 4, 17)
    x0 = W.poisson_state(5, topo["N"], [60, 40, 80, 100, 20], 17)
    n = 64
    w, norm, vmax = _trained(net, [topo] * 8, [x0] * 8, snapshots=16, per_snapshot=1024, steps=600)
    plain = capi.Context(net, [topo] * n, [x0] * n, seed=100)
    plain.run(0.01)
    lrn = capi.Context(net, [topo] * n, [x0] * n, seed=200)
    lrn.mlp_set(w, norm)
    lrn.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=vmax)
    lrn.run(0.01)
    sl = [lrn.stats(r) for r in range(n)]
    assert all(s["status"] == 1 and s["t"] == 0.01 for s in sl)
    props = sum(s["learned_steps"] + s["learned_rejects"] for s in sl)
    rej = sum(s["learned_rejects"] for s in sl) / props
    assert rej <= 0.02, rej
    a = np.stack([plain.get_state(r).reshape(5, -1).sum(1) for r in range(n)]).astype(float)
    b = np.stack([lrn.get_state(r).reshape(5, -1).sum(1) for r in range(n)]).astype(float)
    for s in range(5):
        se = math.sqrt(a[:, s].var(ddof=1) / n + b[:, s].var(ddof=1) / n)
        assert abs(a[:, s].mean() - b[:, s].mean()) < 3 * se + 1e-9, (s, a[:, s].mean(), b[:, s].mean(), se)
        sev = math.sqrt(2 * a[:, s].var() ** 2 / (n - 1) + 2 * b[:, s].var() ** 2 / (n - 1))
        assert abs(a[:, s].var(ddof=1) - b[:, s].var(ddof=1)) < 3 * sev + 1e-9, s
        assert stats.ks_2samp(a[:, s], b[:, s]).pvalue > 0.01, s
    plain.close()
    lrn.close()


def _step_ms(ctx, n=20):
    return ctx.profile(n)[0] / n


def test_learned_tau_full_brusselator_mesh():
    """The learned-tau configuration on the 1,048,576-subvolume Brusselator
    mesh: train on records of a plain run, then 60 learned steps against 60
    plain ones from the same state.

    What the trained network gains over the method without it: on records
    of a later state it was not trained on, its log-tau head explains most
    of the variance of log Cao tau'_v (R^2 > 0.5; the same network
    untrained explains none); its
    steps advance the simulated time about as far as Cao's (within 25 %)
    with <= 2 % of the proposals rejected and no negative population.  The
    per-step cost of both modes is printed (the learned step has no Cao tau
    pass, the network runs in the leap pass), and so is the critical head's
    held-out AUC against the trial-leap reject labels: it is about 0.51, no
    better than the untrained network's, because whether one subvolume's
    trial leap breaks the condition is mostly Poisson noise given its local
    features (an sklearn MLP on the same records reaches 0.54), so the head
    is reported, not asserted."""
    wl = W.brusselator_mesh()
    net, topo, x0 = wl["net"], wl["topos"][0], wl["x0"][0]
    assert topo["N"] == 1048576
    w, norm, vmax = _trained(net, topo, x0, snapshots=8, stride=4, per_snapshot=16384, steps=1500, batch=65536)
    plain = capi.Context(net, [topo], [x0], seed=31)
    plain.set_tau_mode(False, theta=1.0)
    plain.arm(np.inf)
    plain.iterate(60)
    sp = plain.stats()
    plain_ms = _step_ms(plain)
    # held-out records of the state the plain run reached
    f, y, lab = plain.records(0)
    take = np.random.default_rng(3).choice(f.shape[0], size=65536, replace=False)
    f, y, lab = f[take], y[take], lab[take]
    ys = learned.standardize_y(y, norm)
    scores = {}
    for name, weights in (("trained", w), ("untrained", learned.init_weights(11))):
        plain.mlp_set(weights, norm)
        out = plain.mlp_forward(f)
        r2 = 1.0 - np.mean((out[:, 0] - ys) ** 2) / np.var(ys)
        pos, neg = out[lab > 0.5, 1], out[lab < 0.5, 1]
        auc = stats.mannwhitneyu(pos, neg).statistic / (pos.size * neg.size)
        scores[name] = (r2, auc)
    lrn = capi.Context(net, [topo], [x0], seed=31)
    lrn.mlp_set(w, norm)
    lrn.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=vmax)
    lrn.arm(np.inf)
    lrn.iterate(60)
    sl = lrn.stats()
    learned_ms = _step_ms(lrn)
    x = lrn.get_state().reshape(5, -1)
    props = sl["learned_steps"] + sl["learned_rejects"]
    print("plain", sp, plain_ms, "learned", sl, learned_ms, "held-out (R2, AUC)", scores)
    assert scores["trained"][0] > 0.5 and scores["untrained"][0] < 0.1, scores
    assert sl["learned_rejects"] <= 0.02 * props
    assert sl["learned_steps"] >= 0.9 * sl["steps"]
    assert sl["leapbacks"] <= 2
    assert sl["t"] / sl["steps"] > 0.75 * sp["t"] / sp["steps"]
    assert x.sum() > 0
    assert learned_ms < 1.5 * plain_ms
    plain.close()
    lrn.close()


def test_learned_tau_uneven_brusselator_batch():
    """The learned-tau configuration's 128-realization uneven batch (meshes
    of 8..32 per side): records and training on the batch, then learned
    steps on every realization; rejection <= 5 % over the batch, counts
    valid."""
    wl = W.brusselator_batch()
    assert len(wl["topos"]) == 128
    w, norm, vmax = _trained(wl["net"], wl["topos"], wl["x0"], snapshots=6, stride=4, per_snapshot=512,
                             steps=800, batch=65536)
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], seed=41)
    ctx.mlp_set(w, norm)
    ctx.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=vmax)
    ctx.run(np.inf, step_limit=40)
    st = [ctx.stats(r) for r in range(ctx.n)]
    assert all(s["steps"] == 40 for s in st)
    props = sum(s["learned_steps"] + s["learned_rejects"] for s in st)
    rej = sum(s["learned_rejects"] for s in st) / props
    print("batch rejection", rej)
    assert rej <= 0.05
    assert sum(s["learned_steps"] for s in st) >= 0.9 * 40 * 128
    ctx.close()
