"""Shared builders for the parity tests (small cases the oracle finishes in seconds)."""
from __future__ import annotations

import numpy as np

from pliscuda_amd import workloads as W


def asym_topology(N, seed):
    """A directed random graph (not symmetric) with per-edge weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    src, dst = [], []
    for v in range(N):
        k = rng.integers(1, 5)
        nb = rng.choice(N, size=k, replace=False)
        for u in nb:
            src.append(v)
            dst.append(int(u))  # may include v itself (dropped) or duplicates across rows
    src = np.array(src, dtype=np.int64)
    dst = np.array(dst, dtype=np.int64)
    w = rng.uniform(0.5, 1.5, size=src.size)
    return W.csr_from_edges(N, src, dst, w)


def two_compartment_case(seed=5):
    """ABC network with 2 compartments on a 6x6x2 reflecting grid."""
    net = W.network(["A", "B", "C"],
                    [({"A": 1, "B": 1}, {"C": 1}), ({"C": 1}, {"A": 1, "B": 1}), ({}, {"A": 1})],
                    [[0.01, 0.1, 1.0], [0.02, 0.05, 3.0]], [[0.5, 0.5, 0.1], [0.2, 0.8, 0.0]],
                    n_compartments=2)
    topo = dict(W.cubic_grid(6, 6, 2, periodic=False))
    rng = np.random.Generator(np.random.PCG64(seed))
    topo["comp"] = rng.integers(0, 2, size=topo["N"]).astype(np.int32)
    x0 = W.poisson_state(3, topo["N"], [30, 30, 10], seed)
    return net, topo, x0


def low_count_case(seed=11):
    """Low copy numbers: many critical reactions and hops, critical events fire."""
    net = W.abc_network()
    topo = W.cubic_grid(5, 5, 5, periodic=False)
    x0 = W.poisson_state(3, topo["N"], [6.0, 6.0, 3.0], seed)
    return net, topo, x0


def small_bruss(seed=3):
    net = W.brusselator_network(W.SEED_BRUSS)
    topo = W.random_mesh(8, 8, 4, seed)
    x0 = W.poisson_state(5, topo["N"], [60, 40, 80, 100, 20], seed)
    return net, topo, x0


def third_order_case(seed=13):
    """Every reactant pattern of orders 0..3 in one network."""
    net = W.network(["P", "Q", "R"],
                    [({}, {"P": 1}), ({"P": 1}, {"Q": 1}), ({"P": 2}, {"R": 1}),
                     ({"P": 1, "Q": 1}, {"R": 1}), ({"Q": 3}, {"P": 1}),
                     ({"P": 2, "R": 1}, {"Q": 2}), ({"P": 1, "Q": 1, "R": 1}, {}),
                     ({"R": 1}, {"P": 1, "Q": 1})],
                    [5.0, 0.3, 1e-3, 2e-3, 1e-6, 1e-6, 2e-6, 0.2], [0.4, 0.3, 0.2])
    topo = W.cubic_grid(4, 4, 4, periodic=True)
    x0 = W.poisson_state(3, topo["N"], [80, 60, 40], seed)
    return net, topo, x0
