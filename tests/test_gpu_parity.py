"""GPU vs oracle parity: the HIP kernels must reproduce the CPU oracle bit for bit.

Kernel level: per-subvolume tau', a0c and tau'' (plis_debug_tau).
Trajectory level: states and step statistics after many steps, through the
C ABI, across network variants (1 / <=8 / <=64 species), batches of uneven
realizations, compartments, asymmetric weighted graphs, leap-backs and
critical events.
"""
import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import capi
from pliscuda_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _tau_parity(net, topo, x0, eps=0.03, nc=10, seed=1):
    ctx = capi.Context(net, [topo], [x0], eps=eps, nc=nc, seed=seed)
    tp, a0c, t2 = ctx.debug_tau()
    ctx.close()
    m = ffi.Model(net, topo)
    otp, oa0c, ot2 = ffi.tau_all(m, x0, eps, nc, seed)
    assert np.array_equal(tp.view(np.uint64), otp.view(np.uint64)), "tau' differs"
    assert np.array_equal(a0c.view(np.uint64), oa0c.view(np.uint64)), "a0c differs"
    assert np.array_equal(t2.view(np.uint64), ot2.view(np.uint64)), "tau'' differs"
    return otp, oa0c


def _traj_parity(net, topos, xs, eps=0.03, nc=10, seed=1, t_end=np.inf, steps=60, ids=None,
                 use_graph=True, expect=None):
    ctx = capi.Context(net, topos, xs, eps=eps, nc=nc, seed=seed, ids=ids, use_graph=use_graph)
    ctx.run(t_end, step_limit=steps)
    tot = {"leapbacks": 0, "crit_events": 0, "steps": 0}
    for r, (topo, x0) in enumerate(zip(topos, xs)):
        m = ffi.Model(net, topo)
        x = np.array(x0, dtype=np.uint32).copy()
        st = ffi.run(m, x, eps, nc, seed, t_end, step_limit=steps,
                     realization=(ids[r] if ids is not None else r))
        g = ctx.get_state(r)
        gs = ctx.stats(r)
        assert np.array_equal(g, x), f"realization {r}: state differs"
        for f in ("steps", "leapbacks", "firings", "subvolume_leaps", "crit_events"):
            assert gs[f] == getattr(st, f), f"realization {r}: {f} {gs[f]} != {getattr(st, f)}"
        assert gs["t"] == st.t, f"realization {r}: t differs"
        assert gs["status"] == st.status
        for f in tot:
            tot[f] += gs[f]
    ctx.close()
    if expect:
        for f, lo in expect.items():
            assert tot[f] >= lo, f"case did not exercise {f} ({tot[f]} < {lo})"
    return tot


def test_tau_kernel_abc():
    wl = W.abc_reflecting(realizations=1)
    _tau_parity(wl["net"], wl["topos"][0], wl["x0"][0])


def test_tau_kernel_schlogl():
    wl = W.schlogl(n=16)
    _tau_parity(wl["net"], wl["topos"][0], wl["x0"][0])


def test_tau_kernel_brusselator_mesh():
    net, topo, x0 = H.small_bruss()
    _tau_parity(net, topo, x0)


def test_tau_kernel_low_counts_has_critical():
    net, topo, x0 = H.low_count_case()
    tp, a0c = _tau_parity(net, topo, x0)
    assert (a0c > 0).sum() > 10


def test_tau_kernel_third_order():
    net, topo, x0 = H.third_order_case()
    _tau_parity(net, topo, x0)


def test_tau_kernel_compartments():
    net, topo, x0 = H.two_compartment_case()
    _tau_parity(net, topo, x0)


def test_tau_kernel_asymmetric():
    net = W.abc_network()
    topo = H.asym_topology(50, 3)
    x0 = W.poisson_state(3, 50, [40, 40, 20], 3)
    _tau_parity(net, topo, x0)


def test_traj_abc_two_realizations_to_t_end():
    wl = W.abc_reflecting(n=6, realizations=2)
    _traj_parity(wl["net"], wl["topos"], wl["x0"], t_end=2.0, steps=10**9)


def test_traj_schlogl():
    wl = W.schlogl(n=12)
    _traj_parity(wl["net"], wl["topos"], wl["x0"], steps=150)


def test_traj_schlogl_no_graph():
    wl = W.schlogl(n=8)
    _traj_parity(wl["net"], wl["topos"], wl["x0"], steps=50, use_graph=False)


def test_traj_brusselator_mesh():
    net, topo, x0 = H.small_bruss()
    _traj_parity(net, [topo], [x0], steps=80)


def test_traj_low_counts_critical_events():
    net, topo, x0 = H.low_count_case()
    _traj_parity(net, [topo], [x0], steps=300, expect={"crit_events": 5})


def test_traj_leapbacks():
    # a large epsilon makes oversized leaps, so negative populations occur
    net, topo, x0 = H.low_count_case(seed=23)
    _traj_parity(net, [topo], [x0], eps=0.9, nc=2, steps=200, expect={"leapbacks": 1})


def test_traj_third_order():
    net, topo, x0 = H.third_order_case()
    _traj_parity(net, [topo], [x0], steps=60)


def test_traj_compartments():
    net, topo, x0 = H.two_compartment_case()
    _traj_parity(net, [topo], [x0], steps=100)


def test_traj_asymmetric_weighted():
    net = W.abc_network()
    topo = H.asym_topology(60, 9)
    x0 = W.poisson_state(3, 60, [40, 40, 20], 9)
    _traj_parity(net, [topo], [x0], steps=100)


def test_traj_uneven_batch_ids():
    wl = W.schlogl_batch(realizations=6, lo=3, hi=9)
    ids = [100 + 7 * r for r in range(6)]
    _traj_parity(wl["net"], wl["topos"], wl["x0"], t_end=0.05, steps=10**9, ids=ids)


def _wide_network(S=10, R=20):
    """S species, R first- and second-order channels X_a (+ X_b) -> X_c: more
    than 8 species and 16 reactions, so the general instantiation runs and
    the draws use the second coarse block of their counters (hops of species
    8 and 9, reactions 16..19)."""
    names = [f"X{s}" for s in range(S)]
    reactions, rates = [], []
    for j in range(R):
        a, c = j % S, (j * 3 + 1) % S
        if j % 4 == 3:
            b = (j + 5) % S
            reactions.append(({names[a]: 1, names[b]: 1}, {names[c]: 2}))
            rates.append(2e-4 * (1 + j % 3))
        else:
            reactions.append(({names[a]: 1}, {names[c]: 1}))
            rates.append(0.05 * (1 + j % 5))
    return W.network(names, reactions, rates, [0.4 + 0.1 * (s % 4) for s in range(S)])


def test_traj_wide_network_general_instantiation():
    net = _wide_network()
    topo = W.random_mesh(6, 6, 4, 3)
    x0 = W.poisson_state(10, topo["N"], [300.0 + 40 * s for s in range(10)], 12)
    _tau_parity(net, topo, x0)
    _traj_parity(net, [topo], [x0], steps=40)


def test_batch_equals_single():
    """A realization's trajectory does not depend on the batch it runs in."""
    wl = W.schlogl_batch(realizations=4, lo=4, hi=8)
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], seed=5, ids=[0, 1, 2, 3])
    ctx.run(np.inf, step_limit=40)
    ctx1 = capi.Context(wl["net"], wl["topos"][2:3], wl["x0"][2:3], seed=5, ids=[2])
    ctx1.run(np.inf, step_limit=40)
    assert np.array_equal(ctx.get_state(2), ctx1.get_state(0))
    ctx.close()
    ctx1.close()


def _iterate_in_calls(net, topo, x0, calls, eps, nc, seed):
    ctx = capi.Context(net, [topo], [x0], eps=eps, nc=nc, seed=seed)
    ctx.arm(t_end=float("inf"))
    for n in calls:
        ctx.iterate(n)
    x, st = ctx.get_state(0), ctx.stats(0)
    ctx.close()
    return x, st


@pytest.mark.parametrize("case", ["mesh", "leapbacks"])
def test_iterations_split_into_calls_equal_one_call(case):
    """Every call that runs iterations starts by deciding the step on the
    current counts; a decision already in place (or a leap-back's halved
    tau) must come out the same, so splitting the iterations into calls
    changes nothing -- including calls that end right after a leap-back."""
    if case == "mesh":
        net, topo, x0 = H.small_bruss()
        eps, nc, seed, calls = 0.03, 10, 5, (1, 2, 3, 5, 9)
    else:  # the leap-back case of test_traj_leapbacks
        net, topo, x0 = H.low_count_case(seed=23)
        eps, nc, seed, calls = 0.9, 2, 1, (1,) * 10 + (7, 13, 29, 61, 80)
    n = sum(calls)
    xa, sa = _iterate_in_calls(net, topo, x0, (n,), eps, nc, seed)
    xb, sb = _iterate_in_calls(net, topo, x0, calls, eps, nc, seed)
    assert np.array_equal(xa, xb)
    for f in ("steps", "leapbacks", "firings", "crit_events", "t"):
        assert sa[f] == sb[f], (f, sa[f], sb[f])
    if case == "leapbacks":
        assert sa["leapbacks"] >= 1


def test_launch_shape_does_not_change_results():
    wl = W.schlogl(n=16)
    outs = []
    for wpb, bpc in ((1, 1), (4, 8), (2, 3)):
        ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], seed=3, waves_per_block=wpb,
                           blocks_per_cu=bpc)
        ctx.run(np.inf, step_limit=30)
        outs.append(ctx.get_state(0))
        ctx.close()
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_pure_diffusion_conserves_mass_exactly():
    net = W.network(["X", "Y"], [], np.zeros(0), [1.5, 0.2])
    topo = W.cubic_grid(10, 10, 10, periodic=True)
    x0 = np.zeros(2 * topo["N"], dtype=np.uint32)
    x0[0] = 200000
    x0[topo["N"] + 555] = 3000
    ctx = capi.Context(net, [topo], [x0], seed=1)
    ctx.run(np.inf, step_limit=400)
    x = ctx.get_state(0).reshape(2, -1)
    assert x[0].sum() == 200000 and x[1].sum() == 3000
    assert (x[0] > 0).sum() > 100
    ctx.close()


def test_paused_run_equals_straight_run():
    wl = W.schlogl(n=8)
    a = capi.Context(wl["net"], wl["topos"], wl["x0"], seed=9)
    a.run(0.5)
    b = capi.Context(wl["net"], wl["topos"], wl["x0"], seed=9)
    t = 0.0
    while t < 0.5:
        b.run(0.5, t_pause=t + 0.05)
        t = b.stats()["t"]
    assert np.array_equal(a.get_state(), b.get_state())
    assert a.stats()["steps"] == b.stats()["steps"]
    a.close()
    b.close()


def _violations_match(net, topo, x0, eps, steps=15, seed=9, theta=1.0):
    """Every accepted step adds the subvolumes where the post-leap leap
    condition (propensity form, tolerance theta) broke to the statistics;
    step by step this must be the oracle's count between the same two
    states."""
    ctx = capi.Context(net, [topo], [x0], eps=eps, seed=seed)
    ctx.set_tau_mode(False, theta=theta)
    m = ffi.Model(net, topo)
    prev, before, total = np.array(x0, dtype=np.uint32), 0, 0
    for k in range(1, steps + 1):
        ctx.run(np.inf, step_limit=k)
        x = ctx.get_state()
        v = ctx.stats()["violations"]
        n, _ = ffi.violations(m, prev, x, eps, 10, theta)
        assert v - before == n, (k, v - before, n)
        total += n
        prev, before = x, v
    ctx.close()
    return total


@pytest.mark.parametrize("eps", [0.03, 0.1])
def test_violation_counts_match_oracle(eps):
    net, topo, x0 = H.small_bruss()
    assert _violations_match(net, topo, x0, eps) > 0


def test_violation_counts_match_oracle_asymmetric():
    net, topo, x0 = H.small_bruss()
    atopo = H.asym_topology(topo["N"], 3)
    _violations_match(net, atopo, x0, 0.05)
