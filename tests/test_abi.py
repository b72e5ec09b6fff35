"""The C ABI (include/lima.h): the library loads and exports every declared entry point."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lima.h")
LIB = os.path.join(ROOT, "quantom_amd", "liblima_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lima_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_step_api():
    names = declared_functions()
    for n in ("lima_create", "lima_step", "lima_destroy", "lima_get_state", "lima_set_state",
              "lima_get_forces", "lima_get_energies", "lima_get_neighbors", "lima_rebuild_neighbors",
              "lima_last_error", "lima_version"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(lima_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    from quantom_amd.engine import EXPORTS
    assert set(EXPORTS) == set(declared_functions())


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_loads_and_reports_version():
    from quantom_amd import engine
    L = engine.load_library()
    assert b"gfx950" in L.lima_version()
    n = C.c_int(-1)
    assert L.lima_device_count(C.byref(n)) == 0 and n.value >= 0


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_create_without_device_fails_with_status_not_exit(small_system):
    """With no HIP device the engine reports LIMA_E_HIP instead of exiting (Engine.cuh:47-55)."""
    from quantom_amd import engine
    if engine.device_count() > 0:
        pytest.skip("a device is present")
    from quantom_amd.abi import LIMA_E_HIP, RunParams
    with pytest.raises(engine.LimaStatusError) as e:
        engine.Engine([small_system], RunParams())
    assert e.value.code == LIMA_E_HIP


def test_null_arguments_are_rejected():
    from quantom_amd import engine
    L = engine.load_library()
    assert L.lima_create(None, 1, None, None) == 5       # LIMA_E_ARG
    assert L.lima_step(None, 1) == 5
