"""Pins the oracle's per-subvolume model against closed forms: mass-action
propensities (k prod C(x, r), Gillespie 1977; spec/rate_matrix.md:46-58 for
orders 1-2), Cao's critical set (L_j < n_c), Cao's tau' (eq. 33 with the g_i
table) including diffusion in/out terms, and tau'' = -ln(u) / a0c."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import ffi
from pliscuda_amd import workloads as W

SINGLE = {"N": 1, "row_ptr": np.array([0, 0], np.uint32), "col": np.zeros(0, np.uint32)}


def binom(x, r):
    return Fraction(math.comb(x, r)) if x >= r else Fraction(0)


def one_reaction(lhs, rhs, k, species=("A", "B", "C"), d=0.0):
    return W.network(list(species), [(lhs, rhs)], [k], [d] * len(species))


PATTERNS = [
    ({}, {"A": 1}),
    ({"A": 1}, {}),
    ({"A": 2}, {}),
    ({"A": 1, "B": 1}, {}),
    ({"A": 3}, {}),
    ({"A": 2, "B": 1}, {}),
    ({"A": 1, "B": 1, "C": 1}, {}),
]


@pytest.mark.parametrize("lhs,rhs", PATTERNS)
@pytest.mark.parametrize("x", [(0, 0, 0), (1, 2, 3), (2, 5, 1), (7, 11, 13), (1234, 567, 89)])
def test_propensity_closed_form(lhs, rhs, x):
    k = 0.37
    net = one_reaction(lhs, rhs, k)
    m = ffi.Model(net, SINGLE)
    xs = np.array(x, dtype=np.uint32)
    # n_c huge: every reaction of order >= 1 is critical, so a0c = a_j
    _, a0c, _ = ffi.tau_all(m, xs, 0.03, 10**9, 1)
    exact = Fraction(k)
    for s, r in lhs.items():
        exact *= binom(int(xs["ABC".index(s)]), r)
    if not lhs:
        assert a0c[0] == 0.0  # order 0 is never critical
        return
    assert abs(a0c[0] - float(exact)) <= 4e-16 * float(exact) + 1e-300


@pytest.mark.parametrize("x,crit", [(19, True), (20, False), (29, False), (1, True)])
def test_critical_set_cao_L(x, crit):
    """2A -> 0: L = floor(x / 2); critical iff L < n_c = 10."""
    net = one_reaction({"A": 2}, {}, 1.0)
    m = ffi.Model(net, SINGLE)
    _, a0c, _ = ffi.tau_all(m, np.array([x, 0, 0], np.uint32), 0.03, 10, 1)
    assert (a0c[0] > 0) == (crit and x >= 2)


def tau1(net, x, eps=0.03, nc=10):
    m = ffi.Model(net, SINGLE)
    tp, _, _ = ffi.tau_all(m, np.array(x, np.uint32), eps, nc, 1)
    return tp[0]


def test_tau_first_order_decay():
    net = W.network(["X"], [({"X": 1}, {})], [2.0], [0.0])
    # mu = -2000, sigma2 = 2000, g = 1, m = eps x = 30
    assert tau1(net, [1000]) == pytest.approx(min(30 / 2000, 900 / 2000), rel=1e-15)


def test_tau_second_order_same_species():
    net = W.network(["X"], [({"X": 2}, {})], [0.01], [0.0])
    x = 500
    a = 0.01 * x * (x - 1) / 2
    g = 2 + 1 / (x - 1)
    m = max(0.03 * x / g, 1.0)
    assert tau1(net, [x]) == pytest.approx(min(m / (2 * a), m * m / (4 * a)), rel=1e-14)


def test_tau_third_order_g():
    net = W.network(["X"], [({"X": 3}, {})], [1e-6], [0.0])
    x = 800
    a = 1e-6 * math.comb(x, 3)
    g = 3 + 1 / (x - 1) + 2 / (x - 2)
    m = max(0.03 * x / g, 1.0)
    assert tau1(net, [x]) == pytest.approx(min(m / (3 * a), m * m / (9 * a)), rel=1e-14)


def test_tau_mixed_second_order_g_is_2():
    net = W.network(["A", "B"], [({"A": 1, "B": 1}, {})], [1e-3], [0.0, 0.0])
    xa, xb = 400, 900
    a = 1e-3 * xa * xb
    ta = min(max(0.03 * xa / 2, 1) / a, max(0.03 * xa / 2, 1) ** 2 / a)
    tb = min(max(0.03 * xb / 2, 1) / a, max(0.03 * xb / 2, 1) ** 2 / a)
    assert tau1(net, [xa, xb]) == pytest.approx(min(ta, tb), rel=1e-14)


def test_tau_two_node_diffusion_in_and_out():
    """Hops are first-order channels: out-hops nu=-1, in-hops from a
    non-critical neighbour nu=+1 (a sign or a missing term changes tau')."""
    net = W.network(["X"], [], np.zeros(0), [0.5])
    topo = {"N": 2, "row_ptr": np.array([0, 1, 2], np.uint32), "col": np.array([1, 0], np.uint32),
            "w": np.array([1.5, 1.5], np.float32)}
    m = ffi.Model(net, topo)
    tp, _, _ = ffi.tau_all(m, np.array([100, 300], np.uint32), 0.03, 10, 1)
    # node 0: mu = 0.75*(300-100) = 150, s2 = 0.75*400 = 300, m = 3
    assert tp[0] == pytest.approx(min(3 / 150, 9 / 300), rel=1e-15)
    # node 1: mu = -150, s2 = 300, m = 9
    assert tp[1] == pytest.approx(min(9 / 150, 81 / 300), rel=1e-15)
    # a critical neighbour (x < n_c) contributes no in-hop term
    tp, a0c, _ = ffi.tau_all(m, np.array([100, 5], np.uint32), 0.03, 10, 1)
    assert tp[0] == pytest.approx(min(3 / 75, 9 / 75), rel=1e-15)
    assert a0c[1] == pytest.approx(0.75 * 5, rel=1e-15) and a0c[0] == 0.0


def test_tau2_is_exponential_clock_of_critical_propensity():
    net = W.network(["X"], [({"X": 1}, {})], [0.7], [0.0])
    m = ffi.Model(net, SINGLE)
    for step in range(5):
        _, a0c, t2 = ffi.tau_all(m, np.array([4], np.uint32), 0.03, 10, 99, realization=3, step=step)
        w = ffi.philox([0, 0xFFFFFFFF, step, (3 << 16) | (1 << 7)], [99, 0])[0]
        u = (w + 0.5) * 2.0 ** -32
        assert a0c[0] == pytest.approx(2.8)
        assert t2[0] == pytest.approx(-math.log(u) / 2.8, rel=1e-15)


def test_absorbing_state_stops():
    net = W.network(["X"], [({"X": 1}, {})], [1.0], [0.0])
    m = ffi.Model(net, SINGLE)
    x = np.array([0], np.uint32)
    st = ffi.run(m, x, 0.03, 10, 1, 5.0)
    assert st.status == 2 and st.t == 0.0 and st.steps == 0


def test_pure_diffusion_conserves_mass_and_clamps_t_end():
    net = W.network(["X", "Y"], [], np.zeros(0), [2.0, 0.3])
    topo = W.cubic_grid(6, 6, 6, periodic=False)
    x = np.zeros(2 * topo["N"], np.uint32)
    x[5] = 50000
    x[topo["N"] + 100] = 777
    m = ffi.Model(net, topo)
    st = ffi.run(m, x, 0.03, 10, 3, 0.7)
    assert x[: topo["N"]].sum() == 50000 and x[topo["N"]:].sum() == 777
    assert st.t == 0.7 and st.status == 1


def test_leap_back_keeps_populations_valid():
    net = W.abc_network()
    topo = W.cubic_grid(5, 5, 5, periodic=False)
    x = W.poisson_state(3, topo["N"], [6.0, 6.0, 3.0], 17)
    m = ffi.Model(net, topo)
    st = ffi.run(m, x, 0.9, 2, 1, np.inf, step_limit=400)
    assert st.leapbacks >= 1
    assert x.max() < 10**6  # a negative count would have wrapped around


def _coarse_fine_word(c0, c1, purpose, w, key, step=0):
    """The leap draw's 32-bit word: top byte = byte w & 15 of Philox block
    w >> 4 (bytes 0..3 of word 0, then words 1..3), low 24 bits = word w & 3
    of block 64 + (w >> 2); realization 0, attempt 0."""
    o = ffi.philox([c0, c1, step, (purpose << 7) | (w >> 4)], key)
    hb = (o[(w >> 2) & 3] >> (8 * (w & 3))) & 0xFF
    f = ffi.philox([c0, c1, step, (purpose << 7) | (64 + (w >> 2))], key)
    return (hb << 24) | (f[w & 3] & 0xFFFFFF)


def _inversion(lam, word):
    """Poisson(lam) by sequential-search inversion of u = (word + 1/2) 2^-32,
    the CDF summed in exact rationals from a long-double e^-lam."""
    u = Fraction(2 * word + 1, 2 ** 33)
    p = Fraction(str(np.exp(np.longdouble(-lam))))
    F, k, lamf = p, 0, Fraction(lam)
    while u > F:
        k += 1
        p = p * lamf / k
        F += p
    return k


@pytest.mark.parametrize("seed", [1, 7, 23])
def test_leap_hop_draws_by_hand(seed):
    """Two subvolumes, one diffusing species, one step: the hops 0 -> 1 and
    1 -> 0 are the inversion of the words built by hand from Philox (purpose
    4, counter (0, 1, step 0), draws w = 0 and 1), with the means tau d x w
    and tau from the oracle's own tau' (checked against its closed form)."""
    net = W.network(["X"], [], np.zeros(0), [0.5])
    topo = {"N": 2, "row_ptr": np.array([0, 1, 2], np.uint32), "col": np.array([1, 0], np.uint32),
            "w": np.array([1.5, 1.5], np.float32)}
    m = ffi.Model(net, topo)
    x0 = np.array([4000, 1000], np.uint32)
    eps = 0.001
    tp, a0c, _ = ffi.tau_all(m, x0, eps, 10, seed)
    # mu_0 = 0.75 (1000 - 4000), s2_0 = 0.75 * 5000, bound eps x_0 = 4
    assert tp[0] == pytest.approx(min(4 / 2250, 4 ** 2 / 3750), rel=1e-15)
    tau = float(min(tp))
    assert a0c.max() == 0.0
    key = [seed & 0xFFFFFFFF, seed >> 32]
    lam01, lam10 = tau * ((0.5 * 4000.0) * 1.5), tau * ((0.5 * 1000.0) * 1.5)
    k01 = _inversion(lam01, _coarse_fine_word(0, 1, 4, 0, key))
    k10 = _inversion(lam10, _coarse_fine_word(0, 1, 4, 1, key))
    assert 0 < lam10 < lam01 < 10.0
    x = x0.copy()
    st = ffi.run(m, x, eps, 10, seed, np.inf, step_limit=1)
    assert st.steps == 1 and st.leapbacks == 0
    assert list(x) == [4000 - k01 + k10, 1000 + k01 - k10], (k01, k10)
    assert st.firings == k01 + k10


@pytest.mark.parametrize("seed", [2, 5])
def test_leap_reaction_draws_by_hand(seed):
    """One subvolume, 18 first-order channels X -> Y_j (so the draws of
    reactions 16 and 17 come from the second coarse block): each Y_j after one
    step is the inversion of its word built by hand (purpose 3, counter
    (0, 0xFFFFFFFF, step 0), draw w = j), and X lost their sum."""
    R = 18
    species = ["X"] + [f"Y{j}" for j in range(R)]
    reactions = [({"X": 1}, {f"Y{j}": 1}) for j in range(R)]
    rates = [0.01 * (j + 1) for j in range(R)]
    net = W.network(species, reactions, rates, [0.0] * (R + 1))
    m = ffi.Model(net, SINGLE)
    x0 = np.zeros(R + 1, np.uint32)
    x0[0] = 500
    eps = 0.05
    tp, _, _ = ffi.tau_all(m, x0, eps, 10, seed)
    tau = float(tp[0])
    key = [seed & 0xFFFFFFFF, seed >> 32]
    x = x0.copy()
    st = ffi.run(m, x, eps, 10, seed, np.inf, step_limit=1)
    assert st.steps == 1
    want = [_inversion(tau * (rates[j] * 500.0), _coarse_fine_word(0, 0xFFFFFFFF, 3, j, key)) for j in range(R)]
    assert list(x[1:]) == want
    assert x[0] == 500 - sum(want) and sum(want[16:]) > 0
