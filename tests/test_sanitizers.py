"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle (oracle/lima_oracle.cpp)
and the system-ingest parser (quantom_amd/csrc/ingest.cpp) are rebuilt with
-fsanitize=address,undefined (oracle/Makefile, target `sanitize`) and the CPU tests that drive
them through their ctypes bindings run again against those builds in a child process.  Any
sanitizer report ends the child with a non-zero status (UBSan: halt_on_error)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SUITES = ["tests/test_oracle_closed_forms.py", "tests/test_oracle_composition.py", "tests/test_oracle_neighbors.py",
          "tests/test_ingest.py", "tests/test_thermostat.py", "tests/test_reference_layout.py"]


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_and_ingest_under_asan_and_ubsan(tmp_path):
    asan, stdcxx = _runtime("libasan.so"), _runtime("libstdc++.so")
    if asan is None or stdcxx is None:
        pytest.skip("the compiler's ASan runtime is not installed")
    san = str(tmp_path / "san")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j4", "sanitize", f"SAN_DIR={san}"], check=True,
                   capture_output=True)
    env = dict(os.environ)
    # the ASan runtime first; libstdc++ loaded at start-up so that ASan intercepts C++ throws
    env.update(LD_PRELOAD=f"{asan} {stdcxx}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", LIMA_ORACLE_DIR=san,
               LIMA_INGEST_LIB=os.path.join(san, "liblima_ingest.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", *SUITES, "-m", "not gpu", "-x", "-q", "-s",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert " passed" in out
