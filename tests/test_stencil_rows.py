"""The bounded-direction pass evaluates closure rows in a uniform 5-term form
(weights and packed offsets per row kind).  This host-only check compiles
tests/native/stencil_rows_check.cpp against the kernel headers and requires
the 5-term form to equal the row-by-row stencils bit for bit on all 11 row
kinds (interior, 5 left, 5 right rows of src/dfbydx.f:116-228 and
src/d2fdx2.f:116-260), and likewise the mixed-derivative rows of
src/d2fdxy.f:74-391."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_row_form_equals_switch_form(tmp_path):
    exe = tmp_path / "stencil_rows_check"
    src = os.path.join(HERE, "native", "stencil_rows_check.cpp")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "senga-versiona_amd", "csrc"), src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(" ok") == 22
