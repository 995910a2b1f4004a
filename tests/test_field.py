"""Host checks of the field header's fused products and rotations (the
same functions the kernels run), built with hipcc as host code."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_dot_products_and_rotations_are_exact(tmp_path):
    src = tmp_path / "field_dot_test.hip"
    shutil.copy(os.path.join(ROOT, "tests", "cpp", "field_dot_test.cpp"), src)
    exe = str(tmp_path / "field_dot_test")
    cc = HIPCC if os.path.exists(HIPCC) else "hipcc"
    subprocess.run([cc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "tests", "cpp"),
                    str(src), "-o", exe], check=True, capture_output=True, timeout=600)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
