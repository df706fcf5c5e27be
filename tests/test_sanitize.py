"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize/Makefile, gcc/g++ only, no GPU code): every entry point of
the CPU oracle on small grids of each boundary kind, and the runtime's
host-side launch tables (fivonagi_amd/csrc/fv_tables.cpp) checked against
what they must satisfy -- the reference fold's subset against a cell-by-cell
simulation of the fold, work items covering every cell once with the right
CFL lane masks, the adjoint's items and their boundary constraints."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc and g++")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return os.path.join(HERE, "_build")


@pytest.mark.parametrize("prog", ["oracle_check", "host_tables_check"])
def test_sanitized_host_code(built, prog):
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
