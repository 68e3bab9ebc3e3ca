"""Checkpoint / resume through the C ABI: a context restored from a
checkpoint continues bit for bit like the one that wrote it (counts, clock,
step statistics), for a single grid with leap-backs and critical events and
for an uneven batch, also when restored into a context that ran on; a
checkpoint of another seed or epsilon is refused.  Also the per-kernel
timing entry point."""
import numpy as np
import pytest

from pliscuda_amd import capi
from pliscuda_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu

STAT_KEYS = ("t", "steps", "leapbacks", "firings", "subvolume_leaps", "crit_events", "status")


def _resume_matches(net, topos, xs, eps, seed, first, total, t_end=np.inf):
    a = capi.Context(net, topos, xs, eps=eps, seed=seed)
    a.run(t_end, step_limit=first)
    blob = a.checkpoint()
    at_ckpt = [(a.stats(r), a.get_state(r)) for r in range(len(topos))]
    a.run(t_end, step_limit=total)
    b = capi.Context(net, topos, [np.zeros_like(np.asarray(x)) for x in xs], eps=eps, seed=seed)
    b.restore(blob)
    for r, (st, x) in enumerate(at_ckpt):
        assert np.array_equal(b.get_state(r), x)
        assert all(b.stats(r)[k] == st[k] for k in STAT_KEYS)
    b.run(t_end, step_limit=total)
    for r in range(len(topos)):
        assert np.array_equal(a.get_state(r), b.get_state(r)), f"realization {r}: state differs"
        sa, sb = a.stats(r), b.stats(r)
        for k in STAT_KEYS:
            assert sa[k] == sb[k], (r, k, sa[k], sb[k])
    a.close()
    b.close()
    return blob


def test_resume_low_counts_with_critical_events():
    net, topo, x0 = H.low_count_case()
    _resume_matches(net, [topo], [x0], 0.03, 17, 40, 120)


def test_resume_brusselator_mesh():
    net, topo, x0 = H.small_bruss()
    _resume_matches(net, [topo], [x0], 0.03, 5, 25, 60)


def test_resume_uneven_batch_to_t_end():
    wl = W.abc_reflecting(n=6, realizations=5)
    _resume_matches(wl["net"], wl["topos"], wl["x0"], 0.05, 3, 30, 10**6, t_end=2.0)


def test_restore_refuses_other_seed_or_eps():
    net, topo, x0 = H.small_bruss()
    a = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    a.run(np.inf, step_limit=3)
    blob = a.checkpoint()
    a.close()
    for kw in ({"eps": 0.03, "seed": 6}, {"eps": 0.05, "seed": 5}):
        b = capi.Context(net, [topo], [x0], **kw)
        with pytest.raises(capi.PlisError):
            b.restore(blob)
        b.close()
    b = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    with pytest.raises(capi.PlisError):
        b.restore(blob[:100])
    b.close()


def test_restore_refuses_a_model_of_the_same_size():
    """A checkpoint carries a hash of the network, every topology, the edge
    weights and the compartments: one changed rate constant, one rewired edge
    or one changed weight is refused, not restored silently."""
    net, topo, x0 = H.small_bruss()
    a = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    a.run(np.inf, step_limit=3)
    blob = a.checkpoint()
    a.close()
    net2 = dict(net)
    net2["k"] = np.array(net["k"], dtype=np.float64).copy()
    net2["k"][3] *= 1.5
    topo_w = dict(topo)
    topo_w["w"] = np.array(topo["w"], dtype=np.float32).copy()
    # edge 0 (0 -> u) and its reverse u -> 0: the graph stays symmetric
    u = int(topo["col"][0])
    rev = int(topo["row_ptr"][u]) + int(np.searchsorted(topo["col"][topo["row_ptr"][u]:topo["row_ptr"][u + 1]], 0))
    topo_w["w"][[0, rev]] = 1.25
    for n_, t_ in ((net2, topo), (net, topo_w)):
        b = capi.Context(n_, [t_], [x0], eps=0.03, seed=5)
        with pytest.raises(capi.PlisError, match="different network"):
            b.restore(blob)
        b.close()
    b = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    b.restore(blob)  # the same model restores
    b.close()


def test_restore_into_a_context_that_ran_on():
    """Restoring into the context that wrote the checkpoint, after it ran
    further (its leap balance buffers hold the later steps' values), gives
    the same continuation as the first time."""
    net, topo, x0 = H.small_bruss()
    a = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    a.run(np.inf, step_limit=20)
    blob = a.checkpoint()
    a.run(np.inf, step_limit=45)
    first, st1 = a.get_state(), a.stats()
    a.run(np.inf, step_limit=70)
    a.restore(blob)
    a.run(np.inf, step_limit=45)
    assert np.array_equal(a.get_state(), first)
    assert all(a.stats()[k] == st1[k] for k in STAT_KEYS)
    a.close()


def test_profile_kernels_splits_the_leap_pass():
    net, topo, x0 = H.small_bruss()
    a = capi.Context(net, [topo], [x0], eps=0.03, seed=5)
    a.arm(t_end=float("inf"))
    p = a.profile_kernels(5)
    assert len(p) == 8 and all(v >= 0.0 for v in p)
    assert p[5] > 0.0 and p[7] > 0.0  # edge-parallel: draws and finish kernels ran
    assert p[5] + p[6] + p[7] <= p[3] * 1.5 + 0.05
    a.close()
