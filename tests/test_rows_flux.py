"""The flux form of the velocity rows (what the FP32 smoother evaluates)
equals the coefficient form slot by slot: a host program built from the
library's own row header checks S and the diagonal on random data, and that
a planted pair swap is caught (tests/native/rows_flux_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not found")


def test_flux_form_equals_coefficient_form(tmp_path):
    exe = str(tmp_path / "rows_flux_check")
    src = os.path.join(HERE, "native", "rows_flux_check.cpp")
    subprocess.check_call([_hipcc(), "-O2", "-std=c++17", "-o", exe, src], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
