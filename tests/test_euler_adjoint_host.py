"""Hand-written Euler flux adjoints (bosss_amd/csrc/euler_adjoint.hpp) against
forward-mode duals through the forward flux code, compiled and run on the host."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_euler_adjoints_match_dual_jacobians(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "euler_adjoint_check.cpp")
    exe = str(tmp_path / "euler_adjoint_check")
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip", src, "-o", exe],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
