"""Post-processing of simulation snapshots: the reference's state2vtk.py and
octave-scripts (parsestate.m, parseserie.m, countmols.m) restated, plus
ensemble statistics over realizations.

Snapshots are the `sim<timestamp>_[r<k>_]<t>.dat` files the driver writes
(format: PLISCuda/src/driver.cu:435-447 -- subvolumes=, species=, time=, a
blank line, one line of counts per subvolume).  A state is an (N, S) array of
counts, row v = subvolume v.

Subvolume v of an nx x ny x nz grid sits at x = v mod nx, y = (v div nx) mod
ny, z = v div (nx ny): x varies fastest, as in the topologies this package
builds and in countmols.m's ind2sub (octave-scripts/countmols.m:22).
state2vtk.py orders z fastest instead (state2vtk.py:9-13); pass
order="zyx" to the functions here to follow it.

Usage:
  python -m pliscuda_amd.post vtk <snapshot.dat> nx ny nz [--grid] [--seed S] [--order zyx]
  python -m pliscuda_amd.post stats <dir-or-snapshots...>
"""
import glob
import os
import re
import sys

import numpy as np

_NAME = re.compile(r"sim(\d+)_(?:r(\d+)_)?([-+0-9.eE]+)\.dat$")


def parse_snapshot(src):
    """(state (N, S) uint32, time) from a snapshot file path or its text
    (parsestate.m).  Raises ValueError on a malformed header or body."""
    text = open(src).read() if os.path.exists(src) else src
    lines = text.split("\n")
    head = {}
    for i, key in enumerate(("subvolumes", "species", "time")):
        k, _, v = lines[i].partition("=")
        if k.strip() != key:
            raise ValueError(f"line {i + 1}: expected '{key}='")
        head[key] = v.strip()
    N, S, t = int(head["subvolumes"]), int(head["species"]), float(head["time"])
    rows = [ln.split() for ln in lines[4:4 + N]]
    if len(rows) != N or any(len(r) != S for r in rows):
        raise ValueError(f"expected {N} lines of {S} counts")
    return np.array(rows, dtype=np.uint64).astype(np.uint32).reshape(N, S), t


def snapshot_files(path, realization=None):
    """Snapshot files under a directory (or the given list), ordered by the
    time stored in each file (parseserie.m sorts by time, not by name);
    realization=k keeps the files of realization k of a batch run."""
    files = sorted(glob.glob(os.path.join(path, "sim*.dat"))) if isinstance(path, str) else list(path)
    out = []
    for f in files:
        m = _NAME.search(os.path.basename(f))
        if m is None:
            continue
        r = int(m.group(2)) if m.group(2) is not None else None
        if realization is not None and r != realization:
            continue
        out.append(f)
    return sorted(out, key=lambda f: parse_snapshot(f)[1])


def series(path, realization=None):
    """(times (T,), states (T, N, S)) of a run's snapshots (parseserie.m,
    parseseriegrp.m: every species at once)."""
    times, states = [], []
    for f in snapshot_files(path, realization):
        x, t = parse_snapshot(f)
        times.append(t)
        states.append(x)
    if not states:
        raise ValueError("no snapshots found")
    return np.array(times), np.stack(states)


def coords(N, dims, order="xyz"):
    """(x, y, z) integer coordinates of subvolumes 0..N-1 on a dims grid."""
    nx, ny, nz = (int(d) for d in dims)
    if nx * ny * nz != N:
        raise ValueError(f"grid {nx}x{ny}x{nz} does not hold {N} subvolumes")
    v = np.arange(N)
    if order == "xyz":  # x fastest
        return v % nx, (v // nx) % ny, v // (nx * ny)
    if order == "zyx":  # z fastest (state2vtk.py)
        return v // (nz * ny), (v // nz) % ny, v % nz
    raise ValueError("order is 'xyz' or 'zyx'")


def count_mols(state, dims, low, high, order="xyz"):
    """Molecules of each species in the box low..high (corners included),
    countmols.m."""
    state = np.asarray(state)
    x, y, z = coords(state.shape[0], dims, order)
    inside = ((x >= low[0]) & (x <= high[0]) & (y >= low[1]) & (y <= high[1]) & (z >= low[2]) &
              (z <= high[2]))
    return state[inside].astype(np.int64).sum(axis=0)


def ensemble(states):
    """Mean and sample variance of every (subvolume, species) count over
    realizations: states (M, N, S) -> (mean (N, S), var (N, S))."""
    a = np.asarray(states, dtype=np.float64)
    if a.shape[0] < 2:
        raise ValueError("need at least two realizations")
    return a.mean(axis=0), a.var(axis=0, ddof=1)


def write_vtk_points(state, dims, path, seed=0, order="xyz"):
    """One VTK point per molecule at a uniform position inside its
    subvolume's unit cube, with the species index as point scalar
    (state2vtk.py; seeded, so the file is reproducible).  Returns the number
    of points."""
    state = np.asarray(state).astype(np.int64)
    N, S = state.shape
    x, y, z = coords(N, dims, order)
    rng = np.random.default_rng(seed)
    pts, spec = [], []
    for s in range(S):  # all points of species 0, then species 1, ...
        reps = state[:, s]
        cell = np.repeat(np.arange(N), reps)
        base = np.stack([x[cell], y[cell], z[cell]], axis=1).astype(np.float64)
        pts.append(base + rng.random(base.shape))
        spec.append(np.full(len(cell), s))
    pts = np.concatenate(pts) if pts else np.zeros((0, 3))
    spec = np.concatenate(spec) if spec else np.zeros(0, dtype=int)
    n = len(pts)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nmolecules\nASCII\nDATASET POLYDATA\n")
        f.write(f"POINTS {n} float\n")
        np.savetxt(f, pts, fmt="%.6f")
        f.write(f"VERTICES {n} {2 * n}\n")
        np.savetxt(f, np.stack([np.ones(n, dtype=np.int64), np.arange(n)], axis=1), fmt="%d")
        f.write(f"POINT_DATA {n}\nSCALARS species int 1\nLOOKUP_TABLE default\n")
        np.savetxt(f, spec, fmt="%d")
    return n


def write_vtk_grid(state, dims, path, names=None, time=None):
    """The counts as cell data of an nx x ny x nz structured grid, one
    array per species (compact: one number per subvolume and species).
    VTK orders cells x fastest, which is this package's subvolume order."""
    state = np.asarray(state)
    N, S = state.shape
    nx, ny, nz = (int(d) for d in dims)
    if nx * ny * nz != N:
        raise ValueError(f"grid {nx}x{ny}x{nz} does not hold {N} subvolumes")
    names = names or [f"species_{s}" for s in range(S)]
    with open(path, "w") as f:
        title = "counts" + (f" t={time}" if time is not None else "")
        f.write(f"# vtk DataFile Version 2.0\n{title}\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx + 1} {ny + 1} {nz + 1}\nORIGIN 0 0 0\nSPACING 1 1 1\n")
        f.write(f"CELL_DATA {N}\n")
        for s in range(S):
            f.write(f"SCALARS {names[s]} unsigned_int 1\nLOOKUP_TABLE default\n")
            np.savetxt(f, state[:, s], fmt="%d")


def _main(argv):
    if len(argv) >= 5 and argv[0] == "vtk":
        state, t = parse_snapshot(argv[1])
        dims = [int(v) for v in argv[2:5]]
        seed = int(argv[argv.index("--seed") + 1]) if "--seed" in argv else 0
        order = argv[argv.index("--order") + 1] if "--order" in argv else "xyz"
        out = argv[1][:-4] + ".vtk" if argv[1].endswith(".dat") else argv[1] + ".vtk"
        if "--grid" in argv:
            write_vtk_grid(state if order == "xyz" else state[_reorder(len(state), dims)], dims, out, time=t)
            print(f"{out}: {state.shape[0]} cells, {state.shape[1]} species, t = {t}")
        else:
            n = write_vtk_points(state, dims, out, seed=seed, order=order)
            print(f"{out}: {n} molecules, t = {t}")
        return 0
    if len(argv) >= 2 and argv[0] == "stats":
        files = snapshot_files(argv[1]) if len(argv) == 2 and os.path.isdir(argv[1]) else argv[1:]
        for f in files:
            x, t = parse_snapshot(f)
            tot = x.astype(np.int64).sum(axis=0)
            print(f"{os.path.basename(f)}  t={t:g}  " + " ".join(str(int(c)) for c in tot))
        return 0
    print(__doc__)
    return 1


def _reorder(N, dims):
    """Subvolume ids of a z-fastest numbering listed in x-fastest order."""
    nx, ny, nz = (int(d) for d in dims)
    v = np.arange(N)
    x, y, z = v % nx, (v // nx) % ny, v // (nx * ny)
    return (x * ny + y) * nz + z


if __name__ == "__main__":
    sys.exit(_main(sys.argv[1:]))
