"""Runtime checks: one HIP runtime per process (loading the library before torch must leave
torch's GPU usable: PyTorch-ROCm bundles its own libamdhip64 and the loader binds the library
to it), and the device ceilings the library measures."""
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


def test_measured_peaks_are_near_the_published_ceilings():
    """bosss_hip_measure_peaks: the FP64 / FP32 MFMA and FP64 FMA rates it measures lie between
    half and 110 % of the published MI355X figures (78.6, 157.3, 78.6 TFLOP/s)."""
    import ctypes
    from bosss_amd._lib import lib
    from bosss_amd.operator import Context
    print("case measure_peaks", flush=True)
    ctx = Context.get(0)
    v = [ctypes.c_double() for _ in range(3)]
    rc = lib().bosss_hip_measure_peaks(ctx.h, *[ctypes.byref(x) for x in v])
    assert rc == 0, lib().bosss_hip_peaks_last_error().decode()
    for got, pub in zip((x.value for x in v), (78.6, 157.3, 78.6)):
        print("  measured %.1f TFLOP/s of published %.1f" % (got, pub), flush=True)
        assert 0.5 * pub < got < 1.1 * pub
