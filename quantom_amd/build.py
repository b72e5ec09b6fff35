"""Builds the gfx950 shared library quantom_amd/liblima_hip.so with hipcc.

Run as `python -m quantom_amd.build`.  The library is built in-tree so that it
travels with the source tree to the machine that runs it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liblima_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "engine.hip")]
HEADERS = [os.path.join(HERE, "csrc", h) for h in ("device_common.h", "bonded.h", "kernels.h")] + [
    os.path.join(ROOT, "include", "lima.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay disabled (v_pk_*_f32 are never emitted)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-result",
]


def build(verbose: bool = True) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if verbose and warn:
        print("\n".join(warn[:20]))
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
