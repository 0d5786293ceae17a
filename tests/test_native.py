"""Host-side checks of helper arithmetic shared by every kernel (no GPU)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_fastdiv_matches_integer_division(tmp_path):
    src = os.path.join(ROOT, "tests", "native", "fastdiv_check.cpp")
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", src, "-o", exe,
                    "-I/opt/rocm/include"], check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
