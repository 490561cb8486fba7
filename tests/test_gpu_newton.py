"""Implicit-step Newton-Krylov on the GPU (FP64 and FP32-inner defect correction) and the
adjoint gradient of the integrated pressure w.r.t. per-cell energy sources.

Checks: both drivers reach ||F|| <= 1e-10; the converged state also satisfies the
step equation when its residual is evaluated by the CPU oracle; the mixed and FP64
solutions agree; the adjoint gradient matches central finite differences (h = 1e-6).
"""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_EULER_HLLC

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(n=6, p=2, form="euler_hllc"):
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    from bosss_amd.workloads import euler_state
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op64 = DifferentialOperator(g, form, p)
    op32 = DifferentialOperator(g, form, p, precision="fp32")
    Un = torch.from_numpy(euler_state(g, op64.N, seed=7)).cuda()
    dt = 0.5 / n / (2 * p + 1)
    return g, op64, op32, Un, dt


def test_newton_fp64_and_mixed_reach_1e10():
    import torch
    from bosss_amd.newton import ImplicitStep
    say("case newton fp64/mixed")
    g, op64, op32, Un, dt = setup()
    st = ImplicitStep(op64, Un, dt, op32=op32)
    U64 = st.solve_fp64(Un, tol=1e-10)
    s64 = dict(st.stats)
    Um = st.solve_mixed(Un, tol=1e-10)
    sm = dict(st.stats)
    say("  fp64 :", s64)
    say("  mixed:", sm)
    assert s64["history"][-1] <= 1e-10 and sm["history"][-1] <= 1e-10
    assert float(torch.linalg.vector_norm(U64 - Um)) <= 1e-9 * dt * 10
    # independent check of the converged step with the CPU oracle's residual
    ref = RefModel(3, g.ncells, 2, 6, periodic=[1, 1, 1])
    U = Um.cpu().numpy(); Unh = Un.cpu().numpy()
    F = (U - Unh) / dt + ref.evaluate(OP_EULER_HLLC, params_vector(gamma=1.4), U)
    say("  oracle ||F(U_mixed)|| = %.3e" % np.linalg.norm(F))
    assert np.linalg.norm(F) <= 2e-10


def test_newton_with_cell_block_jacobi_preconditioner():
    """Right preconditioning by the inverted diagonal cell blocks of I/dt + J(U_n) (assembled
    by the coloured exact JVPs, inverted by the batched LU): same converged state, fewer
    Krylov iterations; FP64 and the FP32-inner defect correction."""
    import torch
    from bosss_amd.blocks import BlockJacobi
    from bosss_amd.newton import ImplicitStep
    say("case newton block-Jacobi preconditioner")
    g, op64, op32, Un, dt = setup(n=4)
    st = ImplicitStep(op64, Un, dt, op32=op32)
    U0 = st.solve_fp64(Un, tol=1e-10)
    s0 = dict(st.stats)
    ev = op64.GetMatrixBuilder()
    B = ev.new_matrix()
    op64.assemble_dev(Un, B)
    Nt, S = op64.nvars * op64.N, ev.S
    B.view(g.J, S, Nt, Nt)[:, 0] += torch.eye(Nt, dtype=B.dtype, device=B.device) / dt
    bj = BlockJacobi(op64, B)
    bj.check()
    U1 = st.solve_fp64(Un, tol=1e-10, precond=bj.apply_inverse_diagonal)
    s1 = dict(st.stats)

    def precond32(v, z):
        z64 = torch.empty(v.shape, dtype=torch.float64, device=v.device)
        bj.apply_inverse_diagonal(v.double(), z64)
        z.copy_(z64)

    Um = st.solve_mixed(Un, tol=1e-10, precond=precond32)
    sm = dict(st.stats)
    say("  plain      :", s0["jvps"], "JVPs", s0["history"][-1])
    say("  block-Jac  :", s1["jvps"], "JVPs", s1["history"][-1])
    say("  mixed + bJ :", sm["jvps"], "JVPs", sm["history"][-1])
    assert s1["history"][-1] <= 1e-10 and sm["history"][-1] <= 1e-10
    assert s1["jvps"] < s0["jvps"]
    assert float(torch.linalg.vector_norm(U1 - U0)) <= 1e-9 * dt * 10
    assert float(torch.linalg.vector_norm(Um - U0)) <= 1e-9 * dt * 10


def test_adjoint_gradient_matches_finite_differences():
    import torch
    from bosss_amd.newton import ImplicitStep, pressure_functional
    say("case adjoint gradient")
    g, op64, op32, Un, dt = setup(n=4)
    V, N = op64.nvars, op64.N
    S = torch.zeros_like(Un)
    st = ImplicitStep(op64, Un, dt, S=S)
    U = st.solve_fp64(Un, tol=1e-13, inner_tol=1e-8)
    Jv, dJ = pressure_functional(op64, U)
    lam = st.adjoint(U, dJ, tol=1e-12)
    grad = lam.view(g.J, V, N)[:, V - 1, 0]           # d J / d s_j, s_j = energy source on mode 0
    h = 1e-6
    for j in (0, 17, 40):
        vals = []
        for sgn in (1, -1):
            S2 = torch.zeros_like(Un)
            S2.view(g.J, V, N)[j, V - 1, 0] = sgn * h
            st2 = ImplicitStep(op64, Un, dt, S=S2)
            U2 = st2.solve_fp64(U, tol=1e-13, inner_tol=1e-8)
            vals.append(pressure_functional(op64, U2)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        say("  cell %d: adjoint %.10e  FD %.10e" % (j, float(grad[j]), fd))
        assert abs(float(grad[j]) - fd) <= 1e-6 * max(abs(fd), 1e-3)
