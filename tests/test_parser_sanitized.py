"""The host input layer under AddressSanitizer and UBSan (the role of the
reference's cuda-memcheck / initcheck jobs, PLISCuda/test/run.sh:79-120, for
the part of this build that runs on the host): the six reference cases parse
cleanly, and corrupted variants of their files (truncations, non-numeric and
out-of-range tokens, dropped and duplicated lines, reordered configuration
keys) are rejected with a message -- never a crash, an out-of-bounds access
or undefined behaviour."""
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = os.path.join(HERE, "golden", "ref_cases")
HOST = os.path.join(ROOT, "pliscuda_amd", "csrc", "host")
NAMES = ["diff-41", "diff-64", "diff-5125", "react-9", "react-64", "chemotaxis"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=98")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("san") / "parse_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", HOST, "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "sanitize", "parse_check.cpp"), os.path.join(HOST, "plis_inputs.cpp"),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(out)


def _run(harness, d):
    r = subprocess.run([harness, "conf.txt"], cwd=d, capture_output=True, text=True, env=ENV, timeout=60)
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text, text[-3000:]
    assert r.returncode in (0, 2), (r.returncode, text[-2000:])
    return r.returncode, r.stdout


def _copy(name, dst):
    src = os.path.join(CASES, name)
    os.makedirs(dst, exist_ok=True)
    for f in os.listdir(src):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
    return sorted(f for f in os.listdir(dst) if f.endswith(".txt"))


@pytest.mark.parametrize("name", NAMES)
def test_reference_cases_parse_clean(harness, name, tmp_path):
    _copy(name, tmp_path)
    rc, out = _run(harness, tmp_path)
    assert rc == 0 and out.startswith("parsed"), out


def _mutations(text, rng):
    lines = text.split("\n")
    toks = [(i, j) for i, l in enumerate(lines) for j, t in enumerate(l.split(" ")) if t.strip()]
    out = []
    for frac in (0.0, 0.1, 0.5, 0.9, 0.999):
        out.append(text[: int(len(text) * frac)])
    for bad in ("x", "-7", "99999999999999999999999", "", "1e400", ":", "4294967296", "0.5"):
        if not toks:
            break
        i, j = toks[rng.randrange(len(toks))]
        l2 = lines[:]
        parts = l2[i].split(" ")
        parts[j] = bad
        l2[i] = " ".join(parts)
        out.append("\n".join(l2))
    if len(lines) > 3:
        k = rng.randrange(len(lines))
        out.append("\n".join(lines[:k] + lines[k + 1:]))
        out.append("\n".join(lines[:k] + [lines[k]] * 2 + lines[k + 1:]))
    return out


@pytest.mark.parametrize("name", NAMES)
def test_corrupted_inputs_rejected_without_faults(harness, name, tmp_path):
    rng = random.Random(sum(map(ord, name)))
    files = _copy(name, tmp_path / "orig")
    n_run = 0
    for f in files:
        text = open(tmp_path / "orig" / f).read()
        for m, mutated in enumerate(_mutations(text, rng)):
            d = tmp_path / f"{f}_{m}"
            _copy(name, d)
            (d / f).write_text(mutated)
            _run(harness, d)
            n_run += 1
    assert n_run >= 40
