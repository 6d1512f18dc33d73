"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs (CPU only).

tests/sanitize/Makefile builds the CPU oracle and the host half of liblbm355
(lbm355_host.cpp, lbm355_pool.cpp: label setup, open-boundary extrusion,
brick index, partner tables, descriptor validation) with
-fsanitize=address,undefined and -fno-sanitize-recover; the drivers step
every collision model / storage / Carreau-Yasuda coupling of the oracle and
call the device-free entry points on channels, junctions, odd sizes and
random masks.  Any out-of-bounds access, leak or undefined behaviour ends the
driver with a nonzero status and a sanitizer report.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"
KOBJ = HERE.parents[1] / "nonnewlbm_amd" / "_build" / "lbm355_kernels.o"


@pytest.fixture(scope="module")
def built():
    if not KOBJ.exists():
        pytest.skip("the library's kernel object is not built (python -m nonnewlbm_amd.build)")
    r = subprocess.run(["make", "-s", "-C", str(HERE), "-j4"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return HERE / "_build"


@pytest.mark.parametrize("driver", ["oracle_san", "labels_san"])
def test_sanitized_run(built, driver):
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(built / driver)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{driver}: ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr
