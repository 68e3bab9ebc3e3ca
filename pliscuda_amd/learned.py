"""Learned tau proposal: records from plain runs, standardisation, GPU
training of the 12-64-64-2 network and the learned stepping mode.

The network maps 12 local features of a subvolume (log1p counts of the first
five species, log1p of their consuming propensities, highest active reaction
order, log of the previous global tau) to a standardised log tau proposal and
a critical-flag logit.  Training data are records the simulator generates in
plain runs: (features, log Cao tau'_v, accept/reject of a trial leap of v at
tau'_v under the post-leap leap condition).  The log-tau head regresses Cao's
tau', the critical head the reject label.  The learned step uses
tau = min_v tau_v, tau_v = kappa exp(proposal_v), halved where the critical
logit is positive, with the proposals computed inside the previous leap pass;
it is judged after the leap by the leap condition (no negative population and
no more breaking subvolumes than the plain method's rate allows) and falls
back to Cao's tau for that step otherwise.
"""
from __future__ import annotations

import numpy as np

from . import capi

N_PARAMS = capi.MLP_PARAMS
FEATURES = ["log1p x0", "log1p x1", "log1p x2", "log1p x3", "log1p x4",
            "log1p A0", "log1p A1", "log1p A2", "log1p A3", "log1p A4",
            "max active order", "log prev tau"]


def init_weights(seed=0):
    """Glorot-uniform weights, zero biases; rows 12..15 of W1 are padding."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = np.zeros(N_PARAMS, dtype=np.float32)

    def glorot(fan_in, fan_out):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-lim, lim, size=(fan_in, fan_out)).astype(np.float32)

    w1 = np.zeros((16, 64), np.float32)
    w1[:12] = glorot(12, 64)
    o = 0
    w[o:o + 1024] = w1.ravel()
    o += 1024 + 64
    w[o:o + 4096] = glorot(64, 64).ravel()
    o += 4096 + 64
    w[o:o + 128] = glorot(64, 2).ravel()
    return w


def fit_norm(feats, ylog):
    mu = feats.mean(0).astype(np.float32)
    sig = feats.std(0).astype(np.float32)
    sig[sig < 1e-3] = 1.0
    return np.concatenate([mu, sig, [ylog.mean(), max(float(ylog.std()), 1e-3)]]).astype(np.float32)


def standardize_y(ylog, norm):
    return ((ylog - norm[24]) / norm[25]).astype(np.float32)


THETA = 1.0  # tolerance of the post-leap leap condition (the judge and the labels)


def collect_records(ctx, snapshots, stride, per_snapshot, seed=0, t_end=np.inf, theta=THETA):
    """Run ctx in plain mode with the leap condition evaluated (its
    statistics calibrate the judge); every `stride` accepted steps take the
    records of every realization and keep a random subset of `per_snapshot`
    rows.  Returns (features, log Cao tau', reject label)."""
    ctx.set_tau_mode(False, theta=theta)
    rng = np.random.Generator(np.random.PCG64(seed))
    F, Y, Cc = [], [], []
    base = [ctx.stats(r)["steps"] for r in range(ctx.n)]
    for k in range(snapshots):
        ctx.run(t_end, step_limit=min(base) + (k + 1) * stride)
        for r in range(ctx.n):
            f, y, c = ctx.records(r)
            take = rng.choice(f.shape[0], size=min(per_snapshot, f.shape[0]), replace=False)
            F.append(f[take])
            Y.append(y[take])
            Cc.append(c[take])
    return np.concatenate(F), np.concatenate(Y), np.concatenate(Cc)


def train(ctx, feats, ylog, crit, steps=2000, batch=65536, lr=1e-3, seed=1, w0=None):
    """Standardise, set initial weights, train on the GPU (log-tau head on
    log Cao tau', critical head on the reject label); returns (norm, losses)."""
    norm = fit_norm(feats, ylog)
    ctx.mlp_set(init_weights(seed) if w0 is None else w0, norm)
    ys = standardize_y(ylog, norm)
    losses = ctx.mlp_train(feats, ys, crit.astype(np.float32), steps, batch, lr, seed)
    return norm, losses


def calibrate_viol_rate(stats_list, N, margin=2.0):
    """Judge threshold from the plain method's own post-leap leap condition:
    margin x the observed fraction of subvolumes breaking it per subvolume-step.
    N is the realizations' subvolume counts (a list, one per entry of
    stats_list) or one count shared by all."""
    Ns = list(N) if np.ndim(N) else [N] * len(stats_list)
    v = sum(s["violations"] for s in stats_list)
    sv = sum(s["steps"] * n for s, n in zip(stats_list, Ns))
    return margin * v / max(1, sv)
