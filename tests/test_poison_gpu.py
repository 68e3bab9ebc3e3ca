"""Initcheck stand-in (the reference's test/run.sh runs cuda-memcheck
--tool initcheck over its cases, PLISCuda/test/run.sh:100-120): every device
buffer a step writes before it reads is filled with sentinels that are
harmless if read but change the trajectory (plis_debug_poison), at creation
and again between runs; the poisoned run must equal the clean one bit for
bit on every path -- the edge-parallel and the per-subvolume leap, critical
events, an uneven batch, and the learned mode."""
import numpy as np
import pytest

from pliscuda_amd import capi, learned
from pliscuda_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _pair_run(net, topos, x0s, chunks, steps_per_chunk, setup=None, seed=13):
    out = []
    for poison in (False, True):
        ctx = capi.Context(net, topos, x0s, seed=seed)
        if setup:
            setup(ctx)
        for k in range(1, chunks + 1):
            if poison:
                ctx.debug_poison()
            ctx.run(np.inf, step_limit=k * steps_per_chunk)
        out.append(([ctx.get_state(r) for r in range(ctx.n)], [ctx.stats(r) for r in range(ctx.n)]))
        ctx.close()
    (xa, sa), (xb, sb) = out
    for r in range(len(xa)):
        assert np.array_equal(xa[r], xb[r]), f"realization {r}: state differs after poisoning"
        for f in ("steps", "leapbacks", "firings", "crit_events", "t", "learned_steps", "learned_rejects"):
            assert sa[r][f] == sb[r][f], (r, f, sa[r][f], sb[r][f])
    return sa


def test_poison_edge_parallel_leap():
    net, topo, x0 = H.small_bruss()
    _pair_run(net, [topo], [x0], 4, 6)


def test_poison_per_subvolume_leap():
    net, topo, x0 = H.small_bruss()
    _pair_run(net, [H.asym_topology(topo["N"], 3)], [x0], 4, 6)


def test_poison_critical_events_and_leapbacks():
    net, topo, x0 = H.low_count_case()
    st = _pair_run(net, [topo], [x0], 3, 40)
    assert st[0]["crit_events"] > 0


def test_poison_uneven_batch():
    wl = W.schlogl_batch(realizations=6, lo=3, hi=9)
    _pair_run(wl["net"], wl["topos"], wl["x0"], 3, 8)


def test_poison_learned_mode():
    net, topo, x0 = H.small_bruss()
    norm = np.concatenate([np.zeros(12), np.ones(12), [-9.0, 1.0]]).astype(np.float32)

    def setup(ctx):
        ctx.mlp_set(learned.init_weights(7), norm)
        ctx.set_tau_mode(True, kappa=1.0, theta=1.0, viol_rate_max=1.0)

    st = _pair_run(net, [topo], [x0], 4, 6, setup=setup)
    assert st[0]["learned_steps"] > 0
