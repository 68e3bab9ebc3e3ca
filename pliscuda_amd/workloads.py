"""Reaction networks, grids and initial states of the standard workloads.

Every generator is a pure function of its seed (numpy PCG64), so the CPU
oracle and the GPU see identical inputs.

Workloads (N = subvolumes, S = species, R = reactions):
  * ``abc_reflecting``  -- 8x8x8 reflecting grid, A+B->C k=0.01, C->A+B k=0.1,
    0->A k=1.0, d = (0.5, 0.5, 0.1), counts Poisson(50); eps 0.03, n_c 10,
    t_end 10, 64 realizations.
  * ``schlogl``         -- 64^3 periodic grid, Schlogl bistable model with
    the buffers folded into the constants (c1*B1 = 0.03 for 2X->3X with
    C(x,2), c2 = 1e-4 for 3X->2X with C(x,3), c3*B2 = 200 for 0->X,
    c4 = 3.5 for X->0), d = 1, X ~ U{80..600}; eps 0.03, t_end 5.
  * ``brusselator_mesh`` -- 128x128x64 random symmetric mesh, degree 4..12,
    per-edge weight U(0.5, 1.5); 5 species (X, Y, A, B, D), 8 reactions with
    constants log-uniform in [1e-3, 10]; counts Poisson(mean_s), mean_s ~
    U[20, 200].
  * ``schlogl_batch``   -- 256 periodic cubes of side U{16..48}, Schlogl.
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# networks: dict(S, R, C, reac[S*R], prod[S*R], k[C*R], d[C*S], names)
# --------------------------------------------------------------------------


def network(species, reactions, k, d, n_compartments=1):
    """reactions: list of (reactants dict, products dict); k: [C][R] or [R]; d: [C][S] or [S]."""
    S, R = len(species), len(reactions)
    idx = {s: i for i, s in enumerate(species)}
    reac = np.zeros((S, R), dtype=np.int32)
    prod = np.zeros((S, R), dtype=np.int32)
    for j, (lhs, rhs) in enumerate(reactions):
        for s, c in lhs.items():
            reac[idx[s], j] = c
        for s, c in rhs.items():
            prod[idx[s], j] = c
    k = np.asarray(k, dtype=np.float64).reshape(n_compartments, R)
    d = np.asarray(d, dtype=np.float64).reshape(n_compartments, S)
    return {"S": S, "R": R, "C": n_compartments, "reac": reac.ravel(), "prod": prod.ravel(),
            "k": k.ravel(), "d": d.ravel(), "species": list(species)}


def abc_network():
    return network(["A", "B", "C"],
                   [({"A": 1, "B": 1}, {"C": 1}), ({"C": 1}, {"A": 1, "B": 1}), ({}, {"A": 1})],
                   [0.01, 0.1, 1.0], [0.5, 0.5, 0.1])


SCHLOGL_K = (0.03, 1e-4, 200.0, 3.5)


def schlogl_network(d=1.0):
    return network(["X"], [({"X": 2}, {"X": 3}), ({"X": 3}, {"X": 2}), ({}, {"X": 1}), ({"X": 1}, {})],
                   list(SCHLOGL_K), [d])


BRUSSELATOR_REACTIONS = [
    ({"A": 1}, {"A": 1, "X": 1}),          # A -> A + X   (source of X, A buffered)
    ({"B": 1, "X": 1}, {"B": 1, "Y": 1}),  # B + X -> B + Y  (X -> Y)
    ({"X": 2, "Y": 1}, {"X": 3}),          # 2X + Y -> 3X
    ({"X": 1}, {"D": 1}),                  # X -> D
    ({"Y": 1}, {"D": 1}),                  # Y -> D
    ({}, {"Y": 1}),                        # 0 -> Y
    ({"D": 1}, {}),                        # D -> 0
    ({"X": 1}, {}),                        # X -> 0
]
BRUSSELATOR_D = (1.0, 1.0, 0.5, 0.5, 0.2)


def brusselator_network(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    k = 10.0 ** rng.uniform(-3.0, 1.0, size=len(BRUSSELATOR_REACTIONS))
    return network(["X", "Y", "A", "B", "D"], BRUSSELATOR_REACTIONS, k, list(BRUSSELATOR_D))


# --------------------------------------------------------------------------
# topologies: dict(N, row_ptr, col, w=None, comp=None)
# --------------------------------------------------------------------------


def csr_from_edges(N, src, dst, w=None):
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    rp = np.zeros(N + 1, dtype=np.uint32)
    np.add.at(rp, src + 1, 1)
    rp = np.cumsum(rp, dtype=np.uint64).astype(np.uint32)
    out = {"N": int(N), "row_ptr": rp, "col": dst.astype(np.uint32)}
    if w is not None:
        out["w"] = np.asarray(w, dtype=np.float32)[order]
    return out


def cubic_grid(nx, ny, nz, periodic):
    """6-neighbour grid, v = x + nx*(y + ny*z) (the linearisation of the
    reference's genTopology.py:30-37)."""
    N = nx * ny * nz
    v = np.arange(N, dtype=np.int64)
    x, y, z = v % nx, (v // nx) % ny, v // (nx * ny)
    src, dst = [], []
    for axis, n in ((0, nx), (1, ny), (2, nz)):
        for step in (-1, 1):
            c = [x, y, z][axis]
            c2 = c + step
            if periodic:
                c2 = c2 % n
                ok = np.ones(N, dtype=bool)
            else:
                ok = (c2 >= 0) & (c2 < n)
            coords = [x, y, z]
            coords[axis] = c2
            u = coords[0] + nx * (coords[1] + ny * coords[2])
            src.append(v[ok])
            dst.append(u[ok])
    return csr_from_edges(N, np.concatenate(src), np.concatenate(dst))


_HALF_MOORE = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1),
               (0, 1, 1), (0, 1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)]


def random_mesh(nx, ny, nz, seed, dmin=4, dmax=12, wlo=0.5, whi=1.5):
    """Random symmetric mesh on a periodic nx*ny*nz lattice: every subvolume
    links to 4..12 of its 26 Moore neighbours, one weight U(wlo, whi) per
    undirected edge.  Built from 26 perfect matchings (13 half-directions x 2
    parities along the first non-zero axis); a pair is accepted where both
    ends are under their drawn target degree, then a repair pass links
    subvolumes still below dmin."""
    assert nx % 2 == 0 and ny % 2 == 0 and nz % 2 == 0
    rng = np.random.Generator(np.random.PCG64(seed))
    N = nx * ny * nz
    v = np.arange(N, dtype=np.int64)
    xyz = np.stack([v % nx, (v // nx) % ny, v // (nx * ny)])
    dims = np.array([nx, ny, nz])
    target = rng.integers(dmin, dmax + 1, size=N)
    deg = np.zeros(N, dtype=np.int64)
    classes = []
    for d in range(13):
        off = np.array(_HALF_MOORE[d])
        axis = int(np.nonzero(off)[0][0])
        for p in range(2):
            s = v[xyz[axis] % 2 == p]
            t_xyz = (xyz[:, s] + off[:, None]) % dims[:, None]
            t = t_xyz[0] + nx * (t_xyz[1] + ny * t_xyz[2])
            classes.append([s, t, np.zeros(s.size, dtype=bool)])
    for pas in range(3):
        for ci in rng.permutation(len(classes)):
            s, t, added = classes[ci]
            if pas == 0:
                ok = (deg[s] < target[s]) & (deg[t] < target[t]) & (rng.random(s.size) < 0.8)
            else:
                ok = ((deg[s] < dmin) | (deg[t] < dmin)) & (deg[s] < dmax) & (deg[t] < dmax)
            ok &= ~added
            added |= ok
            deg[s[ok]] += 1
            deg[t[ok]] += 1
    src = np.concatenate([s[a] for s, t, a in classes])
    dst = np.concatenate([t[a] for s, t, a in classes])
    assert deg.min() >= dmin and deg.max() <= dmax, (deg.min(), deg.max())
    w = rng.uniform(wlo, whi, size=src.size).astype(np.float32)
    return csr_from_edges(N, np.concatenate([src, dst]), np.concatenate([dst, src]),
                          np.concatenate([w, w]))


# --------------------------------------------------------------------------
# initial states (species-major uint32 [S*N])
# --------------------------------------------------------------------------


def poisson_state(S, N, means, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    means = np.broadcast_to(np.asarray(means, dtype=np.float64), (S,))
    return np.concatenate([rng.poisson(means[s], size=N) for s in range(S)]).astype(np.uint32)


def uniform_state(N, lo, hi, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(lo, hi + 1, size=N).astype(np.uint32)


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
SEED_ABC = 0x0A0B0C
SEED_SCHLOGL = 0x5C410
SEED_BRUSS = 0xB2055
SEED_BATCH = 0xBA7C4


def abc_reflecting(n=8, realizations=64, seed=SEED_ABC):
    net = abc_network()
    topo = cubic_grid(n, n, n, periodic=False)
    N = topo["N"]
    x0 = [poisson_state(3, N, 50.0, seed + r) for r in range(realizations)]
    return {"name": f"abc_reflecting_{n}^3", "net": net, "topos": [topo] * realizations, "x0": x0,
            "eps": 0.03, "nc": 10, "t_end": 10.0, "seed": seed}


def schlogl(n=64, seed=SEED_SCHLOGL):
    net = schlogl_network()
    topo = cubic_grid(n, n, n, periodic=True)
    x0 = [uniform_state(topo["N"], 80, 600, seed)]
    return {"name": f"schlogl_{n}^3_periodic", "net": net, "topos": [topo], "x0": x0,
            "eps": 0.03, "nc": 10, "t_end": 5.0, "seed": seed}


def brusselator_mesh(nx=128, ny=128, nz=64, seed=SEED_BRUSS):
    net = brusselator_network(seed)
    topo = random_mesh(nx, ny, nz, seed + 1)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    means = rng.uniform(20.0, 200.0, size=5)
    x0 = [poisson_state(5, topo["N"], means, seed + 3)]
    return {"name": f"brusselator_mesh_{nx}x{ny}x{nz}", "net": net, "topos": [topo], "x0": x0,
            "eps": 0.03, "nc": 10, "t_end": float("inf"), "seed": seed}


def schlogl_batch(realizations=256, lo=16, hi=48, seed=SEED_BATCH):
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = rng.integers(lo, hi + 1, size=realizations)
    net = schlogl_network()
    topos, x0 = [], []
    cache = {}
    for r, s in enumerate(sides):
        s = int(s)
        if s not in cache:
            cache[s] = cubic_grid(s, s, s, periodic=True)
        topos.append(cache[s])
        x0.append(uniform_state(s ** 3, 80, 600, seed + 1 + r))
    return {"name": f"schlogl_batch_{realizations}x[{lo}..{hi}]^3", "net": net, "topos": topos,
            "x0": x0, "eps": 0.03, "nc": 10, "t_end": 5.0, "seed": seed, "sides": sides}


def brusselator_batch(realizations=128, lo=4, hi=16, seed=SEED_BRUSS + 7):
    """Uneven batch of Brusselator random meshes (the learned-tau batch): each
    realization a (2a x 2b x 2c) mesh with a, b, c ~ U{lo..hi}, the headline
    network and Poisson initial counts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = 2 * rng.integers(lo, hi + 1, size=(realizations, 3))
    net = brusselator_network(SEED_BRUSS)
    means = rng.uniform(20.0, 200.0, size=5)
    topos, x0 = [], []
    for r, (nx, ny, nz) in enumerate(dims):
        t = random_mesh(int(nx), int(ny), int(nz), seed + 100 + r)
        topos.append(t)
        x0.append(poisson_state(5, t["N"], means, seed + 1000 + r))
    return {"name": f"brusselator_batch_{realizations}x[{2 * lo}..{2 * hi}]^3", "net": net, "topos": topos,
            "x0": x0, "eps": 0.03, "nc": 10, "t_end": float("inf"), "seed": seed}


def total_edges(topos):
    return int(sum(len(t["col"]) for t in topos))
