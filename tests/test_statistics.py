"""Statistical correctness of the tau-leap step (oracle on the CPU; the GPU
is bit-identical to it -- tests/test_gpu_parity.py and
tests/test_configs_gpu.py -- and tests/test_mlp_gpu.py compares the learned
mode's ensemble statistics with the plain method's on the GPU).

* closed forms: birth-death stationary Poisson law, first-order decay
  (binomial), mean of pure diffusion e^{tQ} x0;
* exact SSA (direct method): ensemble per-species means and variances within
  3 standard errors, two-sample KS p > 0.01;
* the reference's own golden end states (single samples of its curand
  path): where the reference is unbiased (diff-41, bulk sites) they are
  consistent with this ensemble."""
import math
import os

import numpy as np
import pytest
from scipy import linalg, stats

from oracle import ffi
from pliscuda_amd import inputs
from pliscuda_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "ref_cases")
SINGLE = {"N": 1, "row_ptr": np.array([0, 0], np.uint32), "col": np.zeros(0, np.uint32)}


def ensemble(net, topo, x0, n, t_end, eps=0.03, nc=10, seed=11):
    m = ffi.Model(net, topo)
    xs = [np.array(x0, np.uint32).copy() for _ in range(n)]
    st = ffi.run_batch(m, [topo] * n, xs, eps, nc, seed, t_end)
    assert all(s.status == 1 for s in st)
    return np.stack(xs).astype(np.int64)


def mean_var_ok(sample, mu, var, z=3.0):
    n = sample.size
    se_m = math.sqrt(var / n)
    m4 = var * (1 + 3 * var) if var > 0 else 0  # Poisson-like fourth moment bound
    se_v = math.sqrt(max(m4 - var * var, 2 * var * var) / n)
    return abs(sample.mean() - mu) < z * se_m + 1e-12, abs(sample.var(ddof=1) - var) < z * se_v + 1e-12


def test_birth_death_stationary_poisson():
    net = W.network(["X"], [({}, {"X": 1}), ({"X": 1}, {})], [50.0, 1.0], [0.0])
    x = ensemble(net, SINGLE, [0], 600, 12.0, eps=0.01)[:, 0]
    mu = 50.0 * (1 - math.exp(-12.0))
    okm, okv = mean_var_ok(x, mu, mu)
    assert okm and okv, (x.mean(), x.var())
    ks = np.arange(25, 80)
    obs = np.array([(x == k).sum() for k in ks])
    exp = stats.poisson.pmf(ks, mu) * x.size
    keep = exp > 5
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 0.001


def test_first_order_decay_binomial():
    """Explicit tau-leaping biases a linear decay's mean by about
    exp(-k T eps / 2) (1.5 % at eps = 0.03); at eps = 0.002 that is 0.1 %,
    well inside 3 SE."""
    net = W.network(["A", "B"], [({"A": 1}, {"B": 1})], [0.5], [0.0, 0.0])
    x = ensemble(net, SINGLE, [2000, 0], 500, 2.0, eps=0.002)
    p = math.exp(-1.0)
    okm, okv = mean_var_ok(x[:, 0], 2000 * p, 2000 * p * (1 - p))
    assert okm and okv, (x[:, 0].mean(), x[:, 0].var())
    assert (x.sum(1) == 2000).all()


def test_diffusion_mean_matches_matrix_exponential():
    N, D, T = 21, 2.0, 0.6
    topo = W.cubic_grid(N, 1, 1, periodic=False)
    net = W.network(["X"], [], np.zeros(0), [D])
    x0 = np.zeros(N, np.uint32)
    x0[10] = 20000
    x = ensemble(net, topo, x0, 200, T, eps=0.003)
    A = np.zeros((N, N))
    for v in range(N):
        for q in range(topo["row_ptr"][v], topo["row_ptr"][v + 1]):
            A[v, topo["col"][q]] = D
    Q = A.T - np.diag(A.sum(1))
    mean = linalg.expm(T * Q) @ x0.astype(float)
    emp = x.mean(0)
    se = np.sqrt(np.maximum(x.var(0, ddof=1), 1.0) / x.shape[0])
    z = (emp - mean) / se
    assert (np.abs(z) < 4.5).all(), z
    assert stats.chi2.sf((z ** 2).sum(), N) > 1e-3


def test_tau_leap_matches_exact_ssa_ensemble():
    """ABC network on a 4^3 reflecting grid: per-species totals from 96
    tau-leap realizations vs 96 exact direct-method SSA realizations."""
    net = W.abc_network()
    topo = W.cubic_grid(4, 4, 4, periodic=False)
    x0 = W.poisson_state(3, topo["N"], 50.0, 5)
    t_end, n = 2.0, 96
    tl = ensemble(net, topo, x0, n, t_end).reshape(n, 3, -1).sum(2)
    ex = ffi.ssa_batch(ffi.Model(net, topo), x0, 77, n, t_end).astype(np.int64).sum(2)
    for s in range(3):
        a, b = tl[:, s].astype(float), ex[:, s].astype(float)
        se = math.sqrt(a.var(ddof=1) / n + b.var(ddof=1) / n)
        assert abs(a.mean() - b.mean()) < 3 * se, (s, a.mean(), b.mean(), se)
        # variances: 3 SE of the difference (normal approximation)
        sev = math.sqrt(2 * a.var() ** 2 / (n - 1) + 2 * b.var() ** 2 / (n - 1))
        assert abs(a.var(ddof=1) - b.var(ddof=1)) < 3 * sev
        assert stats.ks_2samp(a, b).pvalue > 0.01


def test_reference_diff41_golden_sample_is_consistent():
    """The reference's single end state of diff-41 (golden/endstate.dat,
    produced by its curand path) scored against this ensemble."""
    c = inputs.read_conf(os.path.join(CASES, "diff-41", "conf.txt"))
    lines = open(os.path.join(CASES, "diff-41", "golden", "endstate.dat")).read().split("\n")
    ref = np.array([int(lines[4 + i].split()[0]) for i in range(41)], dtype=float)
    x = ensemble(c["net"], c["topo"], c["x0"], 64, 1.001805, eps=0.05)
    mu, var = x.mean(0), x.var(0, ddof=1)
    keep = var > 1.0
    z2 = (ref[keep] - mu[keep]) ** 2 / (var[keep] * (1 + 1 / 64))
    # chi-square over the sites, with the ensemble's own sampling error
    assert stats.chi2.sf(z2.sum(), keep.sum()) > 1e-3, (z2.sum(), keep.sum())
    assert ref.sum() == 100000 and (x.sum(1) == 100000).all()


def test_reference_react9_golden_totals_are_typical():
    c = inputs.read_conf(os.path.join(CASES, "react-9", "conf.txt"))
    x = ensemble(c["net"], c["topo"], c["x0"], 200, 8.387568, eps=0.05).reshape(200, 3, -1).sum(2)
    golden = [int(v) for v in open(os.path.join(CASES, "react-9", "golden", "pops.txt")).read().split()]
    assert golden == [0, 0, 900]
    assert (x.sum(1) == 900).all()
    # exact A->B->C chain: E[A] = 900 e^{-t}, E[B] = 900 (e^{-t} - e^{-2t})
    t = 8.387568
    assert abs(x[:, 0].mean() - 900 * math.exp(-t)) < 3 * math.sqrt(900 * math.exp(-t) / 200) + 0.05
    frac = np.mean((x[:, 0] == 0) & (x[:, 1] == 0))
    assert frac > 0.2  # the golden outcome is a typical one
