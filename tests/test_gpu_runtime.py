"""One HIP runtime per process: loading the library before torch must leave torch's GPU usable
(PyTorch-ROCm bundles its own libamdhip64; the loader binds the library to it)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_then_torch_share_one_runtime():
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from bosss_amd._lib import lib\n"
            "from bosss_amd.operator import Context\n"
            "lib(); Context.get(0)\n"
            "import torch\n"
            "x = torch.ones(4, device='cuda:0'); torch.cuda.synchronize()\n"
            "print('ok', float(x.sum()))\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ok 4.0" in r.stdout
