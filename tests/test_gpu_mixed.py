"""FP32-on-MFMA operator variants: error against the FP64 oracle, stated per result.

The FP32 kernels take and return float32 on the device (FP32 accumulation in the
MFMA tiles); the host entry points convert at the boundary.  The bar here is the
single-precision level (relative error per field < 2e-5); the measured errors are
printed so that the log states them.
"""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC

pytestmark = pytest.mark.gpu
KIND = {"advdiff": OP_ADVDIFF, "euler_lf": OP_EULER_LF, "euler_hllc": OP_EULER_HLLC}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("form,p", [("advdiff", 3), ("euler_lf", 2), ("euler_hllc", 2), ("euler_hllc", 3)])
def test_fp32_error_against_fp64_oracle(form, p):
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state
    say("case fp32", form, p)
    n = 4
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    nu = -0.01
    op32 = DifferentialOperator(g, form, p, nu=nu, velocity=(1.0, 0.5, 0.25), precision="fp32")
    ref = RefModel(3, (n,) * 3, p, default_quad_order(form, p), periodic=[1, 1, 1])
    prm = params_vector(1.3, nu, (1.0, 0.5, 0.25), 1.4)
    rng = np.random.default_rng(3)
    U = euler_state(g, op32.N, seed=4) if form.startswith("euler") else rng.uniform(-1, 1, op32.ndof)
    dU = rng.uniform(-1, 1, U.size) * 0.01 * np.abs(U).max()
    W = rng.uniform(-1, 1, U.size)
    V, N = op32.nvars, op32.N
    for name, got, exp in (("residual", op32.evaluate(U), ref.evaluate(KIND[form], prm, U)),
                           ("jvp", op32.jvp(U, dU), ref.jvp(KIND[form], prm, U, dU)),
                           ("vjp", op32.vjp(U, W), ref.vjp(KIND[form], prm, U, W))):
        a, b = got.reshape(-1, V, N), exp.reshape(-1, V, N)
        per_field = max(np.abs(a[:, f] - b[:, f]).max() / np.abs(b[:, f]).max() for f in range(V))
        say("  %s fp32 vs fp64 oracle: per-field %.2e" % (name, per_field))
        assert per_field < 2e-5


@pytest.mark.parametrize("form,p", [("advdiff", 3), ("euler_lf", 2), ("euler_hllc", 2), ("euler_hllc", 3)])
def test_fp32_graded_grid_against_fp64_oracle(form, p):
    """The FP32 kernels on a graded grid (cell widths varying 3x, 3 x 4 x 5 cells): residual,
    JVP, VJP and the fused device residual + VJP within the FP32 bar of the uniform case."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state
    say("case fp32 graded", form, p)
    rng = np.random.default_rng(8)
    nodes = []
    for d in range(3):
        w = rng.uniform(1.0, 3.0, 3 + d)
        nodes.append(np.concatenate([[0.0], np.cumsum(w) / w.sum()]))
    g = CartesianGrid(nodes=nodes, periodic=[True] * 3)
    nu = -0.01
    op32 = DifferentialOperator(g, form, p, nu=nu, velocity=(1.0, 0.5, 0.25), precision="fp32")
    ref = RefModel(3, None, p, default_quad_order(form, p), periodic=[1, 1, 1], nodes=nodes)
    prm = params_vector(1.3, nu, (1.0, 0.5, 0.25), 1.4)
    U = euler_state(g, op32.N, seed=5) if form.startswith("euler") else rng.uniform(-1, 1, op32.ndof)
    dU = rng.uniform(-1, 1, U.size) * 0.01 * np.abs(U).max()
    W = rng.uniform(-1, 1, U.size)
    V, N = op32.nvars, op32.N
    Rref, Gref = ref.evaluate(KIND[form], prm, U), ref.vjp(KIND[form], prm, U, W)
    Ut, Wt = torch.from_numpy(U).cuda().float(), torch.from_numpy(W).cuda().float()
    R, G = torch.zeros_like(Ut), torch.zeros_like(Ut)
    op32.evaluate_vjp_dev(Ut, R, Wt, G)
    torch.cuda.synchronize()
    for name, got, exp in (("residual", op32.evaluate(U), Rref), ("jvp", op32.jvp(U, dU), ref.jvp(KIND[form], prm, U, dU)),
                           ("vjp", op32.vjp(U, W), Gref), ("fused residual", R.double().cpu().numpy(), Rref),
                           ("fused vjp", G.double().cpu().numpy(), Gref)):
        a, b = got.reshape(-1, V, N), exp.reshape(-1, V, N)
        per_field = max(np.abs(a[:, f] - b[:, f]).max() / np.abs(b[:, f]).max() for f in range(V))
        say("  %s fp32 vs fp64 oracle: per-field %.2e" % (name, per_field))
        assert per_field < 2e-5


@pytest.mark.parametrize("form,p", [("euler_lf", 2), ("euler_hllc", 2), ("euler_hllc", 3)])
def test_fp32_supersonic_states_against_fp64_oracle(form, p):
    """The FP32 Euler kernels on supersonic, strongly discontinuous states (rest to Mach 3,
    density and pressure varying up to 6.7x between neighbours, every HLLC region on the faces):
    residual, JVP, VJP and the fused device residual + VJP within the FP32 bar."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from test_gpu_parity import euler_state_supersonic
    say("case fp32 supersonic", form, p)
    n = 4
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op32 = DifferentialOperator(g, form, p, velocity=(1.0, 0.5, 0.25), precision="fp32")
    ref = RefModel(3, (n,) * 3, p, default_quad_order(form, p), periodic=[1, 1, 1])
    prm = params_vector(1.3, 1.0, (1.0, 0.5, 0.25), 1.4)
    U, _ = euler_state_supersonic(g.J, 3, op32.N, g.detJ[0], 23)
    rng = np.random.default_rng(31)
    dU = rng.uniform(-1, 1, U.size) * 0.01 * np.abs(U).max()
    W = rng.uniform(-1, 1, U.size)
    V, N = op32.nvars, op32.N
    Rref, Gref = ref.evaluate(KIND[form], prm, U), ref.vjp(KIND[form], prm, U, W)
    Ut, Wt = torch.from_numpy(U).cuda().float(), torch.from_numpy(W).cuda().float()
    R, G = torch.zeros_like(Ut), torch.zeros_like(Ut)
    op32.evaluate_vjp_dev(Ut, R, Wt, G)
    torch.cuda.synchronize()
    for name, got, exp in (("residual", op32.evaluate(U), Rref), ("jvp", op32.jvp(U, dU), ref.jvp(KIND[form], prm, U, dU)),
                           ("vjp", op32.vjp(U, W), Gref), ("fused residual", R.double().cpu().numpy(), Rref),
                           ("fused vjp", G.double().cpu().numpy(), Gref)):
        a, b = got.reshape(-1, V, N), exp.reshape(-1, V, N)
        per_field = max(np.abs(a[:, f] - b[:, f]).max() / np.abs(b[:, f]).max() for f in range(V))
        say("  %s fp32 vs fp64 oracle: per-field %.2e" % (name, per_field))
        assert per_field < 2e-5
