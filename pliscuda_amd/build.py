"""Builds libplis_hip.so (gfx950) and the ``pliscuda`` driver in-tree.

    python -m pliscuda_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libplis_hip.so")
BIN_DIR = os.path.join(HERE, "bin")
DRIVER = os.path.join(BIN_DIR, "pliscuda")

# -ffp-contract=off: the kernels spell out every fused multiply-add, so the
# host oracle can execute the same IEEE operations.  Packed FP32 is disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    srcs = [os.path.join(CSRC, "plis_capi.hip")]
    _run(["hipcc", *HIP_FLAGS, "-shared", "-o", LIB + ".tmp", *srcs])
    os.replace(LIB + ".tmp", LIB)


def build_driver():
    os.makedirs(BIN_DIR, exist_ok=True)
    host = os.path.join(CSRC, "host")
    srcs = [os.path.join(host, f) for f in ("plis_inputs.cpp", "driver.cpp", "main.cpp")]
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", host,
          "-o", DRIVER + ".tmp", *srcs, "-L", HERE, "-lplis_hip", f"-Wl,-rpath,$ORIGIN/.."])
    os.replace(DRIVER + ".tmp", DRIVER)


def build_all():
    build_lib()
    if os.path.exists(os.path.join(CSRC, "host", "main.cpp")):
        build_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
