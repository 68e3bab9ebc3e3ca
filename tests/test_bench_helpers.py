"""The measurement arithmetic of bench.py, on the CPU: compulsory bytes per
kernel and per step against hand counts, the VALU-issue fraction against its
definition, and the one-GPU guard."""
import json
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _topo(N, E, weighted):
    t = {"N": N, "col": np.zeros(E, dtype=np.uint32)}
    if weighted:
        t["w"] = np.ones(E, dtype=np.float32)
    return t


def test_algorithmic_bytes_hand_counts():
    S, N, E = 5, 1000, 8000
    for wt in (False, True):
        t = [_topo(N, E, wt)]
        topo = 4 * (N + 1) + 4 * E + (4 * E if wt else 0)
        assert bench.algorithmic_bytes(S, t, "tau") == 4 * S * N + topo
        assert bench.algorithmic_bytes(S, t, "leap") == 8 * S * N + topo
        fwd = 4 * (N + 1) + 2 * E + (2 * E if wt else 0)
        assert bench.algorithmic_bytes(S, t, "leap_kernel") == 4 * S * N + fwd + 8 * S * N
        assert bench.algorithmic_bytes(S, t, "finish") == 16 * S * N
        assert bench.algorithmic_bytes(S, t, "step") == N * (12 * S + 10) + 2 * topo
    # a batch sums its realizations
    two = [_topo(N, E, True), _topo(2 * N, 3 * E, True)]
    assert bench.algorithmic_bytes(S, two, "step") == sum(bench.algorithmic_bytes(S, [t], "step") for t in two)


def test_valu_issue_fraction(tmp_path, monkeypatch):
    p = tmp_path / "valu.json"
    p.write_text(json.dumps({"w:k": 24_000_000}))
    monkeypatch.setattr(bench, "PMC_VALU", str(p))
    v = bench.valu_issue("w", "k", 0.1)  # 0.1 ms launches
    t_min = 24e6 * 4 / 1024 / 2.4e9
    assert abs(v["min_us"] - t_min * 1e6) < 1e-9 and abs(v["frac"] - t_min / 1e-4) < 1e-12
    assert bench.valu_issue("w", "missing", 0.1) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
