"""Every benchmark configuration at its full size on the GPU, checked against
the oracle (bit for bit where the oracle finishes in seconds) and by the
properties the configuration fixes where it does not.

* A-B-C 8^3 reflecting grid, 64 realizations, to t = 10: every final state,
  clock and statistic bit-identical to the oracle's batch run.
* Schlogl 64^3 periodic (262,144 subvolumes): tau pass bit-exact, 20 steps
  bit-exact, then the run to t = 5 ends exactly at t_end with counts between
  the model's two stable states' ranges.
* Brusselator 128x128x64 random mesh (1,048,576 subvolumes, ~8.0 M directed
  edges): tau pass bit-exact, 10 steps bit-exact; the same mesh batched with
  two small meshes, 8 steps of each bit-exact.
* 256 Schlogl cubes of side 16..48 in one context: 40 steps of the smallest,
  the largest and six more realizations bit-exact against the oracle, then the
  whole batch to t = 5.
"""
import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import capi
from pliscuda_amd import workloads as W

pytestmark = pytest.mark.gpu

SEED = 2024


def _tau_bits(ctx, model, x0, eps, nc, seed):
    tp, a0c, t2 = ctx.debug_tau()
    otp, oa0c, ot2 = ffi.tau_all(model, x0, eps, nc, seed)
    assert np.array_equal(tp.view(np.uint64), otp.view(np.uint64)), "tau' differs"
    assert np.array_equal(a0c.view(np.uint64), oa0c.view(np.uint64)), "a0c differs"
    assert np.array_equal(t2.view(np.uint64), ot2.view(np.uint64)), "tau'' differs"
    return otp, oa0c


def _same_stats(gs, st, what):
    for f in ("steps", "leapbacks", "firings", "subvolume_leaps", "crit_events", "status"):
        assert gs[f] == getattr(st, f), f"{what}: {f} {gs[f]} != {getattr(st, f)}"
    assert gs["t"] == st.t, f"{what}: t {gs['t']!r} != {st.t!r}"


def test_abc_64_realizations_to_t_end():
    wl = W.abc_reflecting()
    assert len(wl["x0"]) == 64 and wl["topos"][0]["N"] == 512
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=SEED)
    ctx.run(wl["t_end"])
    m = ffi.Model(wl["net"], wl["topos"][0])
    xs = [np.array(x, dtype=np.uint32).copy() for x in wl["x0"]]
    sts = ffi.run_batch(m, wl["topos"], xs, wl["eps"], wl["nc"], SEED, wl["t_end"])
    for r in range(64):
        assert np.array_equal(ctx.get_state(r), xs[r]), f"realization {r}: state differs"
        gs = ctx.stats(r)
        _same_stats(gs, sts[r], f"realization {r}")
        assert gs["t"] == wl["t_end"] and gs["status"] == 1
    ctx.close()


def test_schlogl_64cubed():
    wl = W.schlogl()
    topo, x0 = wl["topos"][0], wl["x0"][0]
    assert topo["N"] == 262144
    m = ffi.Model(wl["net"], topo)
    ctx = capi.Context(wl["net"], [topo], [x0], eps=wl["eps"], nc=wl["nc"], seed=SEED)
    _tau_bits(ctx, m, x0, wl["eps"], wl["nc"], SEED)
    ctx.run(wl["t_end"], step_limit=20)
    x = np.array(x0, dtype=np.uint32).copy()
    st = ffi.run(m, x, wl["eps"], wl["nc"], SEED, wl["t_end"], step_limit=20)
    assert np.array_equal(ctx.get_state(), x), "state after 20 steps differs"
    _same_stats(ctx.stats(), st, "20 steps")
    # the rest of the run to t_end = 5 on the GPU alone
    ctx.run(wl["t_end"])
    s = ctx.stats()
    assert s["status"] == 1 and s["t"] == wl["t_end"], s
    xf = ctx.get_state().astype(np.int64)
    # the bistable model settles between its low (~85) and high (~565) states
    assert 40 < xf.mean() < 700, xf.mean()
    assert s["leapbacks"] <= s["steps"] // 100
    ctx.close()


def test_brusselator_mesh_1M():
    wl = W.brusselator_mesh()
    topo, x0 = wl["topos"][0], wl["x0"][0]
    assert topo["N"] == 1048576 and len(topo["col"]) > 7_000_000
    m = ffi.Model(wl["net"], topo)
    ctx = capi.Context(wl["net"], [topo], [x0], eps=wl["eps"], nc=wl["nc"], seed=SEED)
    _, a0c = _tau_bits(ctx, m, x0, wl["eps"], wl["nc"], SEED)
    ctx.run(np.inf, step_limit=10)
    x = np.array(x0, dtype=np.uint32).copy()
    st = ffi.run(m, x, wl["eps"], wl["nc"], SEED, np.inf, step_limit=10)
    assert np.array_equal(ctx.get_state(), x), "state after 10 steps differs"
    _same_stats(ctx.stats(), st, "10 steps")
    assert st.firings > 10 * 300_000
    ctx.close()


def test_mesh_with_small_realizations_in_one_context():
    """The 1,048,576-subvolume mesh (16,384 tiles: the step decision reduces
    its tile partials in 16 chunk blocks, the last of which decides) in one
    context with two small meshes of the same network (one tile each, whose
    chunk blocks but the first find no partials): 8 steps of every
    realization bit-identical to the oracle's single runs."""
    wl = W.brusselator_mesh()
    net = wl["net"]
    small = [W.random_mesh(4, 4, 4, 5), W.random_mesh(8, 4, 2, 6)]
    topos = [wl["topos"][0]] + small
    x0s = [wl["x0"][0]] + [W.poisson_state(5, t["N"], [60, 40, 80, 100, 20], 7 + i) for i, t in enumerate(small)]
    ctx = capi.Context(net, topos, x0s, eps=wl["eps"], nc=wl["nc"], seed=SEED)
    ctx.run(np.inf, step_limit=8)
    for r, (topo, x0) in enumerate(zip(topos, x0s)):
        m = ffi.Model(net, topo)
        x = np.array(x0, dtype=np.uint32).copy()
        st = ffi.run(m, x, wl["eps"], wl["nc"], SEED, np.inf, step_limit=8, realization=r)
        assert np.array_equal(ctx.get_state(r), x), f"realization {r}: state differs"
        _same_stats(ctx.stats(r), st, f"realization {r}")
    ctx.close()


def test_schlogl_batch_256():
    wl = W.schlogl_batch()
    sides = np.asarray(wl["sides"])
    assert len(wl["topos"]) == 256 and sides.min() >= 16 and sides.max() <= 48
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=SEED)
    ctx.run(wl["t_end"], step_limit=40)
    rng = np.random.default_rng(5)
    pick = {int(np.argmin(sides)), int(np.argmax(sides))}
    pick |= set(int(r) for r in rng.choice(256, size=6, replace=False))
    for r in sorted(pick):
        mr = ffi.Model(wl["net"], wl["topos"][r])
        x = np.array(wl["x0"][r], dtype=np.uint32).copy()
        st = ffi.run(mr, x, wl["eps"], wl["nc"], SEED, wl["t_end"], step_limit=40, realization=r)
        assert np.array_equal(ctx.get_state(r), x), f"realization {r} (side {sides[r]}): state differs"
        _same_stats(ctx.stats(r), st, f"realization {r}")
    ctx.run(wl["t_end"])
    for r in range(256):
        s = ctx.stats(r)
        assert s["status"] == 1 and s["t"] == wl["t_end"], (r, s)
    ctx.close()
