"""Additive Schwarz over multi-cell blocks (bosss_amd.blocks.Schwarz, the host mirror of
Schwarz.Solve, Schwarz.cs:985-1080): one sweep against the same sweep computed densely from
the oracle's operator matrix, with and without overlap, with parts above 128 DOF (factored in
global memory) and on a periodic grid of two cells per direction (a neighbour in two slots);
and Schwarz-preconditioned CG on the SIP matrix, which needs fewer iterations than cell
block-Jacobi."""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_ADVDIFF

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense_sweep(A, parts, Nt, x0, b):
    """x0 + sum_p R_p^T A_p^-1 R_p (b - A x0)."""
    r = b - A @ x0
    z = np.zeros_like(x0)
    for p in parts:
        d = (np.asarray(p)[:, None] * Nt + np.arange(Nt)[None, :]).ravel()
        z[d] += np.linalg.solve(A[np.ix_(d, d)], r[d])
    return x0 + z


@pytest.mark.parametrize("form,shape,p,per,agg,overlap", [
    ("sip", (4, 4, 4), 2, False, (2, 2, 2), 0),        # 8 parts of 80 DOF
    ("sip", (4, 4, 4), 2, False, (2, 2, 2), 1),        # grown by one cell layer: up to 270 DOF
    ("advdiff", (4, 4, 4), 3, True, (2, 2, 2), 0),     # 160-DOF parts: global-memory factorisation
    ("advdiff", (2, 4, 4), 2, True, (2, 2, 2), 0),     # the x neighbour occupies two slots of a row
    ("advdiff", (6, 6, 6), 3, True, (3, 3, 3), 0),     # 8 parts of 540 DOF: two launches per column step
])
def test_schwarz_sweep_matches_dense(form, shape, p, per, agg, overlap):
    import torch
    from bosss_amd.blocks import Schwarz, cartesian_parts
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    say("case schwarz", form, shape, p, per, agg, overlap)
    nu = -0.01 if form == "advdiff" else 1.0
    vel = (1.0, 0.5, 0.25)
    g = CartesianGrid(shape, periodic=[per] * 3)
    op = DifferentialOperator(g, form, p, nu=nu, velocity=vel)
    ref = RefModel(3, shape, p, default_quad_order(form, p), periodic=[int(per)] * 3)
    A = ref.dense_matrix(OP_SIP if form == "sip" else OP_ADVDIFF, params_vector(1.3, nu, vel))
    B = op.GetMatrixBuilder().new_matrix()
    op.assemble_dev(torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0"), B)
    parts = cartesian_parts(g, agg)
    sw = Schwarz(op, B, parts, overlap=overlap)
    sw.check()
    say("  part sizes", sorted(set(len(q) * op.N for q in sw.parts)))
    rng = np.random.default_rng(4)
    x0 = rng.uniform(-1, 1, op.ndof)
    b = rng.uniform(-1, 1, op.ndof)
    x = torch.from_numpy(x0.copy()).cuda()
    sw.Solve(x, torch.from_numpy(b).cuda(), sweeps=1)
    want = dense_sweep(A, sw.parts, op.N, x0, b)
    err = np.abs(x.cpu().numpy() - want).max() / np.abs(want).max()
    say("  sweep vs dense: %.2e" % err)
    assert err < 1e-10


def test_schwarz_preconditioned_cg_beats_cell_block_jacobi():
    import torch
    from bosss_amd.blocks import BlockJacobi, Schwarz, cartesian_parts, pcg
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    say("case schwarz pcg")
    g = CartesianGrid((8, 8, 8))
    op = DifferentialOperator(g, "sip", 2)
    B = op.GetMatrixBuilder().new_matrix()
    op.assemble_dev(torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0"), B, alpha=-1.0)   # -SIP is SPD
    b = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, op.ndof)).cuda()
    A = lambda v, out: op.spmv_dev(B, v, out)
    bj = BlockJacobi(op, B)
    _, it_bj, rel_bj = pcg(A, b, bj.apply_inverse_diagonal, tol=1e-10)
    its = {}
    for overlap in (0, 1):
        sw = Schwarz(op, B, cartesian_parts(g, (2, 2, 2)), overlap=overlap)
        sw.check()
        x, its[overlap], rel = pcg(A, b, sw.apply, tol=1e-10)
        r = torch.empty_like(b)
        A(x, r)
        true = float(torch.linalg.vector_norm(r - b) / torch.linalg.vector_norm(b))
        say("  overlap %d: %d iterations (cell block-Jacobi %d), true rel residual %.2e" % (overlap, its[overlap], it_bj, true))
        assert rel <= 1e-10 and true < 1e-9
    assert its[0] < it_bj and its[1] < its[0]
