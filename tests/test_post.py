"""Post-processing (the reference's state2vtk.py and octave-scripts restated):
snapshot parsing round trip, time-ordered series, box counts against a
brute-force count, VTK files parsed back (one point per molecule inside its
own cell; cell data in VTK's x-fastest order), ensemble statistics."""
import os

import numpy as np
import pytest

from pliscuda_amd import inputs, post


def _state(N, S, seed, hi=6):
    return np.random.default_rng(seed).integers(0, hi, size=(N, S)).astype(np.uint32)


def _write(d, name, x, t):
    N, S = x.shape
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(inputs.snapshot_text(x.T.reshape(-1), S, N, t))
    return p


def test_snapshot_round_trip(tmp_path):
    x = _state(24, 3, 1, hi=1000)
    y, t = post.parse_snapshot(_write(tmp_path, "sim1_0.500000.dat", x, 0.5))
    assert np.array_equal(x, y) and t == 0.5
    with pytest.raises(ValueError):
        post.parse_snapshot("subvolumes=2\nspecies=1\nclock=0\n\n1 \n2 \n")


def test_series_sorted_by_time_and_realization(tmp_path):
    xs = [_state(8, 2, s) for s in range(6)]
    # names out of time order; a second realization interleaved
    _write(tmp_path, "sim7_r0_2.000000.dat", xs[2], 2.0)
    _write(tmp_path, "sim7_r0_10.000000.dat", xs[3], 10.0)
    _write(tmp_path, "sim7_r0_1.000000.dat", xs[1], 1.0)
    _write(tmp_path, "sim7_r1_1.000000.dat", xs[4], 1.0)
    _write(tmp_path, "sim7_r1_2.000000.dat", xs[5], 2.0)
    t, s = post.series(str(tmp_path), realization=0)
    assert list(t) == [1.0, 2.0, 10.0]
    assert np.array_equal(s, np.stack([xs[1], xs[2], xs[3]]))
    t, s = post.series(str(tmp_path), realization=1)
    assert list(t) == [1.0, 2.0] and np.array_equal(s[1], xs[5])


@pytest.mark.parametrize("order", ["xyz", "zyx"])
def test_count_mols_matches_brute_force(order):
    dims = (4, 3, 5)
    N = 60
    x = _state(N, 3, 2, hi=50)
    for low, high in (((0, 0, 0), (3, 2, 4)), ((1, 1, 1), (2, 2, 3)), ((3, 0, 4), (3, 0, 4))):
        ref = np.zeros(3, dtype=np.int64)
        for v in range(N):
            if order == "xyz":
                cx, cy, cz = v % 4, (v // 4) % 3, v // 12
            else:
                cz, cy, cx = v % 5, (v // 5) % 3, v // 15
            if all(lo <= c <= hi for c, lo, hi in zip((cx, cy, cz), low, high)):
                ref += x[v]
        assert np.array_equal(post.count_mols(x, dims, low, high, order), ref)


def _read_vtk(path):
    return open(path).read().split("\n")


def test_vtk_points_one_per_molecule_inside_its_cell(tmp_path):
    dims = (3, 2, 2)
    x = _state(12, 2, 3, hi=4)
    p = str(tmp_path / "a.vtk")
    n = post.write_vtk_points(x, dims, p, seed=5)
    assert n == int(x.sum())
    lines = _read_vtk(p)
    assert lines[4] == f"POINTS {n} float"
    pts = np.array([[float(v) for v in ln.split()] for ln in lines[5:5 + n]])
    k = lines.index(f"POINT_DATA {n}")
    spec = np.array([int(v) for v in lines[k + 3:k + 3 + n]])
    # points of species s are listed cell by cell: rebuild the expected cells
    cells = np.concatenate([np.repeat(np.arange(12), x[:, s]) for s in range(2)])
    assert np.array_equal(spec, np.concatenate([np.full(int(x[:, s].sum()), s) for s in range(2)]))
    cx, cy, cz = cells % 3, (cells // 3) % 2, cells // 6
    lo = np.stack([cx, cy, cz], axis=1)
    assert np.all(pts >= lo - 1e-6) and np.all(pts < lo + 1 + 1e-6)
    p2 = str(tmp_path / "b.vtk")
    post.write_vtk_points(x, dims, p2, seed=5)
    assert open(p).read() == open(p2).read()


def test_vtk_grid_cell_data_in_vtk_order(tmp_path):
    dims = (4, 3, 2)
    x = _state(24, 2, 4, hi=100)
    p = str(tmp_path / "g.vtk")
    post.write_vtk_grid(x, dims, p, names=["A", "B"], time=1.5)
    lines = _read_vtk(p)
    assert "DIMENSIONS 5 4 3" in lines and "CELL_DATA 24" in lines
    for s, name in enumerate(("A", "B")):
        k = lines.index(f"SCALARS {name} unsigned_int 1")
        vals = np.array([int(v) for v in lines[k + 2:k + 2 + 24]])
        assert np.array_equal(vals, x[:, s])  # package order is VTK's x-fastest order
    # a z-fastest numbering is put into VTK order
    v = np.arange(24)
    cx, cy, cz = v % 4, (v // 4) % 3, v // 12  # VTK cell v
    zfast = (cx * 3 + cy) * 2 + cz
    assert np.array_equal(post._reorder(24, dims), zfast)


def test_ensemble_statistics():
    rng = np.random.default_rng(6)
    st = rng.poisson(30.0, size=(40, 10, 3)).astype(np.uint32)
    m, v = post.ensemble(st)
    assert np.allclose(m, st.mean(axis=0)) and np.allclose(v, st.astype(float).var(axis=0, ddof=1))
    with pytest.raises(ValueError):
        post.ensemble(st[:1])


def test_cli_vtk(tmp_path, capsys):
    x = _state(8, 2, 7)
    p = _write(tmp_path, "sim1_0.250000.dat", x, 0.25)
    assert post._main(["vtk", p, "2", "2", "2", "--seed", "1"]) == 0
    assert os.path.exists(p[:-4] + ".vtk")
    assert post._main(["vtk", p, "2", "2", "2", "--grid"]) == 0
    assert post._main(["stats", str(tmp_path)]) == 0
    out = capsys.readouterr().out
    assert f"t=0.25  {int(x[:, 0].sum())} {int(x[:, 1].sum())}" in out
