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
