"""The compacted sampler of the edge-parallel leap pass.

phase_b_kernel settles draws with the zero shortcut on the top byte of
their words only, queues the rest per wavefront (at most 256, with their
top bytes) and draws them after the walk, one per lane (fine bits, float
squeeze, double search where it cannot decide); a tile with a mean >= 10 draw or a full queue applies nothing from
its queue and phase_b_redo_kernel draws its open draws with the complete
sampler.  Each undirected edge belongs to the tile of its lower end, which
draws both of its hop directions.  These cases drive each branch on purpose
-- draws queued and drained, queues that overflow on some tiles only, every
tile redone -- predict on the host which tiles the leap pass must hand to
the redo pass (the same draws and the same zero tests, rebuilt from the
oracle's Philox), check the GPU's count against that, and check the
trajectories bit for bit against the oracle; one case runs the
multi-compartment instantiations of the same kernels.
"""
import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import capi
from pliscuda_amd import workloads as W

pytestmark = pytest.mark.gpu

P_REACT, P_HOP = 3, 4
QUEUE = 256
F32 = np.float32


def _tau(net, topo, x0, eps, nc=10, seed=1):
    tp, _, t2 = ffi.tau_all(ffi.Model(net, topo), x0, eps, nc, seed)
    return min(tp.min(), t2.min())


def _binom(x, r):
    if x < r:
        return 0.0
    xd = float(x)
    if r == 1:
        return xd
    if r == 2:
        return (xd * (xd - 1.0)) * 0.5
    return ((xd * (xd - 1.0)) * (xd - 2.0)) / 6.0


def _tune_eps(net, topo, x0, lam_of_tau, lo_t, hi_t):
    """eps whose first-step largest Poisson mean lies in [lo_t, hi_t]."""
    lo, hi = 1e-5, 1.0
    for _ in range(80):
        eps = (lo * hi) ** 0.5
        lam = lam_of_tau(_tau(net, topo, x0, eps))
        if lam < lo_t:
            lo = eps
        elif lam > hi_t:
            hi = eps
        else:
            return eps
    raise AssertionError("no eps found")


def _coarse_byte(o, w):
    """Byte w & 15 of a Philox block: bytes 0..3 of word 0, then words 1..3."""
    return (o[(w >> 2) & 3] >> (8 * (w & 3))) & 0xFF


def _hop_mean_f(tau, d, x):
    """A hop's mean in single precision like the kernel: tau_f ((d_f x_f) w), w = 1."""
    return F32(tau) * (F32(F32(d) * F32(x)) * F32(1.0))


def _hop_coarse_zero_f(tau, d, x, hb):
    """The walk's zero shortcut for a hop, on the top byte of its word and in
    single precision like the kernel: mean l = tau_f ((d_f x_f) w) with
    w = 1, settled iff hb <= (255 - 2^-12) - l (256 + 2^-10)."""
    lf = _hop_mean_f(tau, d, x)
    if not lf > 0:
        return True
    g = F32(F32(255.0 - 2.0 ** -12) - F32(lf * F32(256.0 + 2.0 ** -10)))
    return bool(F32(hb) <= g)


def _react_coarse_zero(lam, hb):
    if not lam > 0.0:
        return True
    return (hb + 1) * 2.0 ** -8 <= (1.0 - lam) - 2.0 ** -50


def _predict_step0(net, topo, x0, eps, seed=1, nc=10):
    """Per tile of the first leap pass: number of draws the zero shortcut
    leaves open (queued) and whether a draw with a mean that may be 10 or
    more occurs (hops: float mean above 9.9999; reactions: double mean >= 10).  Unweighted
    symmetric topology, one compartment, at most 8 species, realization 0,
    step 0, attempt 0."""
    S, R, N = net["S"], net["R"], topo["N"]
    x = x0.reshape(S, N).astype(np.int64)
    d = np.asarray(net["d"], dtype=float)
    k = np.asarray(net["k"], dtype=float)
    reac = np.asarray(net["reac"]).reshape(S, R)
    order = reac.sum(0)
    tau = _tau(net, topo, x0, eps, nc, seed)
    rp = topo["row_ptr"].astype(np.int64)
    col = topo["col"].astype(np.int64)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    ntiles = (N + 63) // 64
    und = np.zeros(ntiles, dtype=np.int64)
    big = np.zeros(ntiles, dtype=bool)

    for v in range(N):
        t = v // 64
        for q in range(rp[v], rp[v + 1]):
            u = col[q]
            if u < v:
                continue  # each undirected edge is drawn by the tile of its lower end
            o = ffi.philox((v, u, 0, P_HOP << 7), key)  # the edge's coarse block
            for s in range(S):
                for src, dirbit in ((v, 0), (u, 1)):
                    xs = x[s, src]
                    if d[s] > 0 and xs >= nc and xs > 0:
                        if not _hop_coarse_zero_f(tau, d[s], xs, _coarse_byte(o, 2 * s + dirbit)):
                            und[t] += 1
                            # the walk defers a tile with a float mean above 9.9999
                            big[t] |= bool(_hop_mean_f(tau, d[s], xs) > F32(9.9999))
        o, oblk = None, -1
        for j in range(R):
            aj = k[j]
            L = -1
            for s in range(S):
                r = int(reac[s, j])
                if r:
                    aj = aj * _binom(x[s, v], r)
                    L = x[s, v] // r if L < 0 else min(L, x[s, v] // r)
            if (order[j] > 0 and L < nc) or not aj > 0.0:
                continue
            if j >> 4 != oblk:
                o, oblk = ffi.philox((v, 0xFFFFFFFF, 0, (P_REACT << 7) | (j >> 4)), key), j >> 4
            if not _react_coarse_zero(tau * aj, _coarse_byte(o, j)):
                und[t] += 1
                big[t] |= tau * aj >= 10.0
    return und, big


def _first_steps_match(net, topo, x0, eps, steps, expect_redo, seed=1):
    ctx = capi.Context(net, [topo], [x0], eps=eps, seed=seed)
    ctx.run(np.inf, step_limit=1)
    redo = ctx.redo_tiles()
    st = ctx.stats()
    assert st["leapbacks"] == 0  # the prediction is for attempt 0
    assert redo == expect_redo, (redo, expect_redo)
    m = ffi.Model(net, topo)
    xo = x0.copy()
    so = ffi.run(m, xo, eps, 10, seed, np.inf, step_limit=1)
    assert np.array_equal(ctx.get_state(), xo) and st["firings"] == so.firings
    ctx.run(np.inf, step_limit=steps)
    xo = x0.copy()
    so = ffi.run(m, xo, eps, 10, seed, np.inf, step_limit=steps)
    st = ctx.stats()
    assert np.array_equal(ctx.get_state(), xo), "state differs"
    for f in ("steps", "leapbacks", "firings", "subvolume_leaps"):
        assert st[f] == getattr(so, f), f
    assert st["t"] == so.t
    ctx.close()


def test_queued_draws_drained_in_tile():
    """Conversion X0 <-> X1 plus diffusion on a 32 x 16 grid with means
    around 0.3: every tile queues some of its ~600 draws (hops and
    reactions), none overflows, and the drained draws give the oracle's
    trajectory."""
    net = W.network(["X0", "X1"], [({"X0": 1}, {"X1": 1}), ({"X1": 1}, {"X0": 1})], [1.0, 1.0],
                    [1.0 / 6.0, 1.0 / 6.0])
    topo = W.cubic_grid(32, 16, 1, periodic=False)
    x0 = W.poisson_state(2, topo["N"], 2000.0, 11)
    eps = _tune_eps(net, topo, x0, lambda tau: tau * float(x0.max()), 0.2, 0.45)
    und, big = _predict_step0(net, topo, x0, eps)
    assert not big.any() and und.min() > 0 and und.max() <= QUEUE, (und, big)
    _first_steps_match(net, topo, x0, eps, 20, expect_redo=0)


def test_queue_overflow_sends_some_tiles_to_redo():
    """Eight diffusing species on a mesh of degree 8-16 (about 6,000 draws per
    tile) with means tuned so that the open draws of some tiles exceed the
    queue: those tiles (and only those) are redone."""
    S = 8
    net = W.network([f"X{s}" for s in range(S)], [({}, {"X0": 1})], [0.0], [1.0] * S)
    topo = W.random_mesh(8, 8, 8, 7, dmin=8, dmax=16, wlo=1.0, whi=1.0)
    x0 = W.poisson_state(S, topo["N"], 2000.0, 3)
    found = None
    for lam_hi in (0.040, 0.044, 0.048, 0.036, 0.052, 0.032, 0.056):
        eps = _tune_eps(net, topo, x0, lambda tau: tau * float(x0.max()), lam_hi * 0.99, lam_hi * 1.01)
        und, big = _predict_step0(net, topo, x0, eps)
        n_over = int((und > QUEUE).sum())
        if not big.any() and 0 < n_over < len(und):
            found = (eps, n_over)
            break
    assert found, und
    _first_steps_match(net, topo, x0, found[0], 10, expect_redo=found[1])


def test_large_means_redo_every_tile():
    """Birth-death around x = 1000 with a coarse eps: means in the hundreds
    (PTRS) in every tile, so the whole leap pass runs in the redo kernel."""
    net = W.network(["X"], [({}, {"X": 1}), ({"X": 1}, {})], [5000.0, 5.0], [0.1])
    topo = W.cubic_grid(8, 8, 8, periodic=True)
    x0 = W.poisson_state(1, topo["N"], 1000.0, 5)
    und, big = _predict_step0(net, topo, x0, 0.3)
    assert big.all()
    _first_steps_match(net, topo, x0, 0.3, 10, expect_redo=len(big))


def test_large_means_with_compartments_redo_every_tile():
    """Two compartments (per-subvolume rate and diffusion tables: the
    multi-compartment instantiations of the leap, redo and finish kernels)
    with means in the hundreds: every tile goes to the redo pass on the
    first step, and the trajectory stays bit-identical to the oracle."""
    net = W.network(["X", "Y"], [({}, {"X": 1}), ({"X": 1}, {}), ({"X": 1}, {"Y": 1}), ({"Y": 1}, {})],
                    [[5000.0, 5.0, 0.5, 1.0], [3000.0, 2.0, 1.0, 0.5]], [[0.1, 0.2], [0.3, 0.05]],
                    n_compartments=2)
    topo = dict(W.cubic_grid(8, 8, 8, periodic=True))
    topo["comp"] = (np.arange(topo["N"]) // 100 % 2).astype(np.int32)
    x0 = W.poisson_state(2, topo["N"], [1000.0, 800.0], 9)
    ctx = capi.Context(net, [topo], [x0], eps=0.3, seed=4)
    ctx.run(np.inf, step_limit=1)
    st = ctx.stats()
    assert st["leapbacks"] == 0 and ctx.redo_tiles() == topo["N"] // 64
    ctx.run(np.inf, step_limit=8)
    m = ffi.Model(net, topo)
    xo = x0.copy()
    so = ffi.run(m, xo, 0.3, 10, 4, np.inf, step_limit=8)
    st = ctx.stats()
    assert np.array_equal(ctx.get_state(), xo), "state differs"
    for f in ("steps", "leapbacks", "firings", "subvolume_leaps"):
        assert st[f] == getattr(so, f), f
    ctx.close()
