"""Input surface: the reference's file formats parsed by the Python reader and
by the C++ driver (pliscuda --parse-only) agree, and malformed inputs fail
with the reference's messages."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from pliscuda_amd import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "ref_cases")
DRIVER = os.path.join(os.path.dirname(HERE), "pliscuda_amd", "bin", "pliscuda")
NAMES = ["diff-41", "diff-64", "diff-5125", "react-9", "react-64", "chemotaxis"]
# sizes listed by the reference's test/tests.txt and the case files
EXPECTED = {"diff-41": (41, 1, 1, 1), "diff-64": (64, 2, 1, 1), "diff-5125": (5125, 1, 1, 1),
            "react-9": (9, 3, 2, 1), "react-64": (64, 2, 1, 1), "chemotaxis": (60, 32, 62, 2)}


@pytest.mark.parametrize("name", NAMES)
def test_reader_sizes(name):
    c = inputs.read_conf(os.path.join(CASES, name, "conf.txt"))
    N, S, R, C = EXPECTED[name]
    assert (c["topo"]["N"], c["net"]["S"], c["net"]["R"], c["net"]["C"]) == (N, S, R, C)
    assert c["x0"].size == S * N


def test_reader_semantics_react9():
    c = inputs.read_conf(os.path.join(CASES, "react-9", "conf.txt"))
    # reactions.txt: "1 0 0 -> 0 1 0", "0 1 0 -> 0 0 1", layout [s*R + j]
    assert list(c["net"]["reac"]) == [1, 0, 0, 1, 0, 0]
    assert list(c["net"]["prod"]) == [0, 0, 1, 0, 0, 1]
    assert list(c["net"]["k"]) == [1.0, 2.0] and list(c["net"]["d"]) == [float(np.float32(0.01))] * 3
    assert c["topo"]["col"][:3].tolist() == [1, 0, 2]  # "0: 1", "1: 0 2"
    assert c["x0"].reshape(3, 9)[0].tolist() == [100] * 9


def test_chemotaxis_compartments():
    c = inputs.read_conf(os.path.join(CASES, "chemotaxis", "conf.txt"))
    comp = c["topo"]["comp"]
    assert comp[0] == 0 and (comp[1:] == 1).all()
    assert c["net"]["k"].size == 2 * 62  # constants2.txt carries one extra value, ignored


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
@pytest.mark.parametrize("name", NAMES)
def test_driver_parse_only_matches_reader(name):
    d = os.path.join(CASES, name)
    out = subprocess.run([DRIVER, "--parse-only", "conf.txt"], cwd=d, capture_output=True, text=True,
                         check=True).stdout
    m = re.search(r"parsed N=(\d+) S=(\d+) R=(\d+) C=(\d+) E=(\d+) molecules=(\d+)", out)
    c = inputs.read_conf(os.path.join(d, "conf.txt"))
    assert m, out
    vals = [int(g) for g in m.groups()]
    assert vals == [c["topo"]["N"], c["net"]["S"], c["net"]["R"], c["net"]["C"], len(c["topo"]["col"]),
                    int(c["x0"].astype(np.int64).sum())]


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
@pytest.mark.parametrize("mutation,msg", [
    (lambda s: s.replace("topology=", "topologyX="), "need a topology file"),
    (lambda s: s.replace("endTime=", "endtime="), "need an endtime value"),
    (lambda s: s.replace("state.txt", "missing.txt"), "file missing.txt not found"),
])
def test_driver_reports_bad_conf(tmp_path, mutation, msg):
    src = os.path.join(CASES, "react-9")
    for f in os.listdir(src):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(src, f), tmp_path / f)
    conf = (tmp_path / "conf.txt").read_text()
    (tmp_path / "conf.txt").write_text(mutation(conf))
    r = subprocess.run([DRIVER, "--parse-only", "conf.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and msg in r.stdout
    with pytest.raises(ValueError, match=msg):
        inputs.read_conf(str(tmp_path / "conf.txt"))


def test_state_rejects_negative_and_wrong_length(tmp_path):
    p = tmp_path / "s.txt"
    p.write_text("subvolumes: 2\nspecies: 2\n\n0: 1 -3\n")
    with pytest.raises(ValueError, match="species count < 0"):
        inputs.read_state(str(p))
    p.write_text("subvolumes: 2\nspecies: 2\n\n0: 1\n")
    with pytest.raises(ValueError, match="wrong length"):
        inputs.read_state(str(p))


def test_topology_rejects_out_of_range(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("subvolumes: 2\n\n0: 1\n1: 2\n")
    with pytest.raises(ValueError, match="sub >= n"):
        inputs.read_topology(str(p))


def test_snapshot_format_matches_golden_layout():
    """The snapshot writer reproduces the reference's layout byte for byte
    (golden/endstate.dat of react-9 re-rendered from its own numbers)."""
    g = open(os.path.join(CASES, "react-9", "golden", "endstate.dat")).read()
    lines = g.split("\n")
    N, S = int(lines[0].split("=")[1]), int(lines[1].split("=")[1])
    t = float(lines[2].split("=")[1])
    x = np.array([[int(v) for v in lines[4 + i].split()] for i in range(N)]).T.ravel()
    assert inputs.snapshot_text(x, S, N, t) == g
