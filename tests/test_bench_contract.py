"""bench.py's contract on the host: more than one GPU is refused with a message before any GPU
work, and the algorithmic flop counts behind roofline.achieved are the dense-contraction
counts of the reference formulation, restated here per configuration from the per-cell /
per-edge rule (2 V N Kv terms per cell, 4 V N Ke terms per edge)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("form,D,n,p,per,expected", [
    # Euler p=3 64^3: Kv = 125, Ke = 25, V = 5, N = 20; volume: values + 3 test gradients,
    # edges: values + test values, 3 x 64^3 periodic faces
    ("euler_hllc", 3, 64, 3, True, 2 * 5 * 20 * 125 * 4 * 64 ** 3 + 4 * 5 * 20 * 25 * 2 * 3 * 64 ** 3),
    # upwind + SIP p=3 64^3: Kv = 64, Ke = 16, V = 1; volume (1 + 3) + 3, edges (1 + 3) + (1 + 3)
    ("advdiff", 3, 64, 3, True, 2 * 20 * 64 * 7 * 64 ** 3 + 4 * 20 * 16 * 8 * 3 * 64 ** 3),
    # Euler p=2 32^3: Kv = 64, Ke = 16, N = 10
    ("euler_lf", 3, 32, 2, True, 2 * 5 * 10 * 64 * 4 * 32 ** 3 + 4 * 5 * 10 * 16 * 2 * 3 * 32 ** 3),
    # SIP p=2 64^2 with Dirichlet boundary: Kv = 9, Ke = 3, N = 6; 2 x 64 x 63 inner + 4 x 64 boundary edges
    ("sip", 2, 64, 2, False, 2 * 6 * 9 * 4 * 64 ** 2 + 4 * 6 * 3 * 6 * (2 * 64 * 63 + 4 * 64)),
])
def test_algorithmic_flops_per_apply(form, D, n, p, per, expected):
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.workloads import contraction_flops
    g = CartesianGrid((n,) * D, periodic=[per] * D)
    assert contraction_flops(form, D, p, g.J, len(g.cells)) == expected


def test_cpu_baseline_records_the_host_layout():
    sys.path.insert(0, ROOT)
    import bench
    info = bench.host_cpu_layout()
    assert info["cpus_usable"] >= 1
    assert int(info.get("cpus", info["cpus_usable"])) >= info["cpus_usable"]
