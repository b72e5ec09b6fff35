"""Drop-in callers: a C++ program written against include/lima_engine.hpp (the
reference's Engine method names, Engine.cuh:109-208) and a C program against
include/lima.h compile, link to the library and get errors as values (no GPU
here, so creating an engine must fail with a status, never exit())."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

LIBDIR = os.path.join(ROOT, "quantom_amd")
LIB = os.path.join(LIBDIR, "liblima_hip.so")

CPP = r"""
#include "lima_engine.hpp"
#include <cstdio>
int main() {
    std::vector<float> x(3, 1.0f);
    int32_t off[1] = {0};
    lima_system s{};
    s.box_len = 5.0f; s.dt = 2e-6; s.thermostat_target = 300.0f;
    s.compound_offset = off; s.n_solvents = 1;
    s.solvent_positions = x.data(); s.solvent_prev_positions = x.data();
    try {
        lima::Engine e(s, lima::default_params());
        e.deviceMaster();
        e.hostMaster();
        auto p = e.parameters(0);
        std::printf("engine %zu\n", p.size());
    } catch (const lima::Error& err) {
        std::printf("error %d\n", err.code);
    }
    return 0;
}
"""

C_SRC = r"""
#include "lima.h"
#include <stdio.h>
int main(void) {
    lima_engine* eng = 0;
    int st = lima_create(0, 1, 0, &eng);
    lima_adjoint_spec spec = {0};
    printf("%d %d %s\n", st, lima_adjoint_run(0, &spec, 0, 0), lima_version());
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("g++") is None, reason="library or g++ missing")
def test_cpp_engine_header_compiles_links_and_reports_errors(tmp_path):
    src = tmp_path / "env.cpp"
    src.write_text(CPP)
    exe = tmp_path / "env"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", LIBDIR, "-llima_hip", f"-Wl,-rpath,{LIBDIR}"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0
    line = out.stdout.strip()
    # no GPU in this container: an error code, not a crash or exit(); with a GPU: an engine
    assert line.startswith("error ") or line == "engine 68", line


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("gcc") is None, reason="library or gcc missing")
def test_c_header_compiles_and_rejects_null_arguments(tmp_path):
    src = tmp_path / "caller.c"
    src.write_text(C_SRC)
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", LIBDIR, "-llima_hip", f"-Wl,-rpath,{LIBDIR}"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    st_create, st_adj, version = out.stdout.split(" ", 2)
    assert st_create == "5" and st_adj == "5"          # LIMA_E_ARG
    assert version.startswith("lima-mi355x")
