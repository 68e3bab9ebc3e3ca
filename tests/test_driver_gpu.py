"""End-to-end: the `pliscuda conf.txt` driver on the reference's six test cases
(PLISCuda/test/*), checked the way the reference's test/run.sh checks them
(population totals from the final sim*.dat against golden/pops.txt) where
the golden is an invariant, against every stoichiometric conservation law
otherwise, and bit-for-bit against the oracle from the same seed."""
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
from scipy.linalg import null_space

from oracle import ffi
from pliscuda_amd import inputs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "ref_cases")
BENCH = os.path.join(HERE, "golden", "ref_benchmarks")
DRIVER = os.path.join(os.path.dirname(HERE), "pliscuda_amd", "bin", "pliscuda")
SEED = 4321


def run_case(name, tmp_path, extra="", root=CASES):
    src = os.path.join(root, name)
    for f in os.listdir(src):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(src, f), tmp_path / f)
    conf = (tmp_path / "conf.txt").read_text().rstrip("\n") + f"\nepsilon=0.05\nnc=10\nseed={SEED}\n{extra}"
    (tmp_path / "conf.txt").write_text(conf)
    print(f"driver case {name}", flush=True)
    r = subprocess.run([DRIVER, "conf.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    for key in ("final simulation time:", "elapsed time:", "simulation steps:", "steps/second:"):
        assert key in r.stdout
    snaps = sorted(glob.glob(str(tmp_path / "sim*.dat")))
    assert snaps
    final = max(snaps, key=lambda p: float(re.search(r"_([0-9.]+)\.dat$", p).group(1)))
    lines = open(final).read().split("\n")
    N, S = int(lines[0].split("=")[1]), int(lines[1].split("=")[1])
    t = float(lines[2].split("=")[1])
    x = np.array([[int(v) for v in lines[4 + i].split()] for i in range(N)], dtype=np.int64)
    return r.stdout, t, x.T.copy(), final, snaps


def conservation_ok(net, x0, x1):
    S, R = net["S"], net["R"]
    nu = (np.asarray(net["prod"]) - np.asarray(net["reac"])).reshape(S, R).astype(float)
    L = null_space(nu.T)  # laws L^T nu = 0
    if L.size == 0:
        return True
    a = L.T @ x0.reshape(S, -1).sum(1).astype(float)
    b = L.T @ x1.reshape(S, -1).sum(1).astype(float)
    return np.allclose(a, b, rtol=0, atol=1e-6 * max(1.0, np.abs(a).max()))


@pytest.mark.parametrize("name", ["diff-41", "diff-64", "react-9", "react-64", "chemotaxis"])
def test_reference_case_matches_oracle(name, tmp_path):
    out, t, x, final, _ = run_case(name, tmp_path)
    c = inputs.read_conf(os.path.join(CASES, name, "conf.txt"))
    assert t == pytest.approx(c["t_end"], abs=1e-6)  # the last step is clamped to endTime
    model = ffi.Model(c["net"], c["topo"])
    xo = c["x0"].copy()
    ffi.run(model, xo, 0.05, 10, SEED, c["t_end"])
    assert np.array_equal(x.ravel(), xo.astype(np.int64)), name
    assert conservation_ok(c["net"], c["x0"].astype(np.int64), x.ravel())
    pops = open(os.path.join(CASES, name, "golden", "pops.txt")).read().split()
    if name.startswith("diff"):
        # pure diffusion: totals are an exact invariant, golden pops must match
        assert [str(v) for v in x.reshape(c["net"]["S"], -1).sum(1)] == pops


def test_reference_case_diff5125_mass_and_front(tmp_path):
    out, t, x, final, _ = run_case("diff-5125", tmp_path)
    assert int(x.sum()) == 100000  # golden/pops.txt
    assert (x > 0).sum() > 200  # the front has spread from subvolume 2562


def test_log_freq_snapshots_do_not_change_the_trajectory(tmp_path):
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _, ta, xa, _, _ = run_case("react-9", a)
    src = os.path.join(CASES, "react-9")
    for f in os.listdir(src):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(src, f), b / f)
    conf = (b / "conf.txt").read_text().replace("logFreq=0", "logFreq=1")
    (b / "conf.txt").write_text(conf.rstrip("\n") + f"\nepsilon=0.05\nnc=10\nseed={SEED}\n")
    r = subprocess.run([DRIVER, "conf.txt"], cwd=b, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0
    snaps = glob.glob(str(b / "sim*.dat"))
    assert len(snaps) >= 6  # one per >= 1 time unit after the previous one, plus the final
    final = max(snaps, key=lambda p: float(re.search(r"_([0-9.]+)\.dat$", p).group(1)))
    lines = open(final).read().split("\n")
    xb = np.array([[int(v) for v in lines[4 + i].split()] for i in range(9)], dtype=np.int64).T
    assert np.array_equal(xa, xb)


def test_checkpoint_restore_continues_the_run(tmp_path):
    """`pliscuda` with checkpoint= stops at endTime 4 and saves the state;
    a second run with restore= continues to endTime 8 and lands where one
    context stepping to 4 and then to 8 lands (the run to 4 clamps its last
    step, so the pair is compared with the same two-leg run through the C
    ABI).  stateOut= writes the final state as a re-readable State file."""
    from pliscuda_amd import capi
    src = os.path.join(CASES, "react-9")
    a, b = tmp_path / "a", tmp_path / "b"
    for d in (a, b):
        d.mkdir()
        for f in os.listdir(src):
            if f.endswith(".txt"):
                shutil.copyfile(os.path.join(src, f), d / f)
    base = (a / "conf.txt").read_text().rstrip("\n").replace("endTime=8", "endTime=4")
    (a / "conf.txt").write_text(base + f"\nepsilon=0.05\nnc=10\nseed={SEED}\ncheckpoint=ck.bin\nstateOut=final.txt\n")
    print("driver react-9 to t=4 with checkpoint", flush=True)
    r = subprocess.run([DRIVER, "conf.txt"], cwd=a, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    conf_b = (b / "conf.txt").read_text().rstrip("\n")
    (b / "conf.txt").write_text(conf_b + f"\nepsilon=0.05\nnc=10\nseed={SEED}\nrestore={a / 'ck.bin'}\n")
    print("driver react-9 restored to t=8", flush=True)
    r = subprocess.run([DRIVER, "conf.txt"], cwd=b, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restored from" in r.stdout
    snaps = glob.glob(str(b / "sim*.dat"))
    final = max(snaps, key=lambda p: float(re.search(r"_([0-9.]+)\.dat$", p).group(1)))
    lines = open(final).read().split("\n")
    N, S = int(lines[0].split("=")[1]), int(lines[1].split("=")[1])
    assert float(lines[2].split("=")[1]) == pytest.approx(8.0)
    x = np.array([[int(v) for v in lines[4 + i].split()] for i in range(N)], dtype=np.uint32).T.ravel()
    c = inputs.read_conf(os.path.join(src, "conf.txt"))
    ctx = capi.Context(c["net"], [c["topo"]], [c["x0"]], eps=0.05, nc=10, seed=SEED)
    ctx.run(4.0)
    x4 = ctx.get_state()
    ctx.run(8.0)
    assert np.array_equal(x, ctx.get_state())
    ctx.close()
    st = inputs.read_state(str(a / "final.txt"))
    assert st["N"] == N and st["S"] == S and np.array_equal(st["x"], x4)
    # a checkpoint of another seed is refused with a non-zero status
    (b / "conf.txt").write_text(conf_b + f"\nepsilon=0.05\nnc=10\nseed={SEED + 1}\nrestore={a / 'ck.bin'}\n")
    r = subprocess.run([DRIVER, "conf.txt"], cwd=b, capture_output=True, text=True, timeout=110)
    assert r.returncode != 0 and "cannot restore" in r.stderr


def test_realizations_with_log_freq_all_reach_end_time(tmp_path):
    """realizations=2 with logFreq: realizations step with their own tau, so
    they cross their snapshot times at different iterations; every one of
    them must still run to endTime, get its own periodic snapshots and end
    where the oracle ends it."""
    src = os.path.join(CASES, "react-9")
    for f in os.listdir(src):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(src, f), tmp_path / f)
    conf = (tmp_path / "conf.txt").read_text().replace("logFreq=0", "logFreq=1")
    (tmp_path / "conf.txt").write_text(conf.rstrip("\n") + f"\nepsilon=0.05\nnc=10\nseed={SEED}\nrealizations=2\n")
    print("driver react-9 x2 realizations with logFreq=1", flush=True)
    r = subprocess.run([DRIVER, "conf.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    c = inputs.read_conf(os.path.join(src, "conf.txt"))
    model = ffi.Model(c["net"], c["topo"])
    for q in range(2):
        snaps = glob.glob(str(tmp_path / f"sim*_r{q}_*.dat"))
        times = sorted(float(re.search(r"_([0-9.]+)\.dat$", p).group(1)) for p in snaps)
        assert len(times) >= 6, (q, times)  # >= one per time unit, plus the final
        assert times[-1] == pytest.approx(c["t_end"], abs=1e-6), (q, times)
        assert all(b - a >= 1.0 - 1e-5 for a, b in zip(times[:-2], times[1:-1])), (q, times)
        final = max(snaps, key=lambda p: float(re.search(r"_([0-9.]+)\.dat$", p).group(1)))
        lines = open(final).read().split("\n")
        N = int(lines[0].split("=")[1])
        x = np.array([[int(v) for v in lines[4 + i].split()] for i in range(N)], dtype=np.int64).T.ravel()
        xo = c["x0"].copy()
        ffi.run(model, xo, 0.05, 10, SEED, c["t_end"], realization=q)
        assert np.array_equal(x, xo.astype(np.int64)), q


@pytest.mark.parametrize("name", ["diff-64", "react-64", "chemotaxis", "diff-5125"])
def test_reference_benchmark_inputs(name, tmp_path):
    """The reference's benchmark inputs (PLISCuda/benchmarks/*, the cases its
    benchmarks/run.sh times) through the driver: the run reaches endTime,
    reports steps/second and subvolume-leaps/second, keeps every
    stoichiometric conservation law, and matches the oracle bit for bit
    (diff-5125, 200 k pure-tau'' steps: exact mass and a spread front)."""
    out, t, x, final, _ = run_case(name, tmp_path, root=BENCH)
    c = inputs.read_conf(os.path.join(BENCH, name, "conf.txt"))
    assert t == pytest.approx(c["t_end"], abs=1e-6)
    for key in ("steps/second:", "subvolume-leaps/second:", "firings:"):
        assert key in out
    assert conservation_ok(c["net"], c["x0"].astype(np.int64), x.ravel())
    if name == "diff-5125":
        assert int(x.sum()) == 100000 and (x > 0).sum() > 200
        return
    model = ffi.Model(c["net"], c["topo"])
    xo = c["x0"].copy()
    ffi.run(model, xo, 0.05, 10, SEED, c["t_end"])
    assert np.array_equal(x.ravel(), xo.astype(np.int64)), name
