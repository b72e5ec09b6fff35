"""Builds the gfx950 shared library quantom_amd/liblima_hip.so with hipcc, and the
host-only ingest library quantom_amd/liblima_ingest.so with g++.

Run as `python -m quantom_amd.build`.  The libraries are built in-tree so that they
travel with the source tree to the machine that runs them.
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

INGEST_LIB = os.path.join(HERE, "liblima_ingest.so")
INGEST_SOURCES = [os.path.join(HERE, "csrc", "ingest.cpp")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# FP32 in source order: the ingest's parameters and coordinates must equal the
# reference's bit for bit, so no FMA contraction.
INGEST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wextra"]
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
    build_ingest(verbose)
    return LIB


def build_ingest(verbose: bool = True) -> str:
    cmd = [CXX, *INGEST_FLAGS, "-o", INGEST_LIB + ".tmp", *INGEST_SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed on the ingest library")
    os.replace(INGEST_LIB + ".tmp", INGEST_LIB)
    return INGEST_LIB


if __name__ == "__main__":
    build()
