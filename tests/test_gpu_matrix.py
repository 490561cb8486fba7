"""Operator-matrix assembly (coloured exact JVP probes) and block-ELL SpMV on the GPU,
against the oracle's exact Jacobian blocks (bref_assemble)."""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_ADVDIFF, OP_EULER_HLLC

pytestmark = pytest.mark.gpu
KIND = {"sip": OP_SIP, "advdiff": OP_ADVDIFF, "euler_hllc": OP_EULER_HLLC}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("form,D,n,p,per", [("sip", 2, 6, 2, False), ("advdiff", 3, 4, 3, True), ("sip", 3, 3, 2, False),
                                            ("euler_hllc", 3, 3, 2, True),
                                            ("sip", 2, 64, 2, False)])   # the 64 x 64 p = 2 Dirichlet Poisson problem
def test_assembled_blocks_and_spmv(form, D, n, p, per):
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state
    say("case assemble", form, D, n, p, per)
    g = CartesianGrid((n,) * D, periodic=[per] * D)
    nu = -0.01 if form == "advdiff" else 1.0
    vel = (1.0, 0.5, 0.25)[:D]
    op = DifferentialOperator(g, form, p, nu=nu, velocity=vel)
    ref = RefModel(D, (n,) * D, p, default_quad_order(form, p), periodic=[int(per)] * D)
    prm = params_vector(1.3, nu, vel, 1.4)
    U = euler_state(g, op.N, seed=3) if form.startswith("euler") else np.zeros(op.ndof)
    Ut = torch.from_numpy(U).cuda()
    ev = op.GetMatrixBuilder()
    B = ev.new_matrix()
    op.assemble_dev(Ut, B)
    x = np.random.default_rng(1).uniform(-1, 1, op.ndof)
    xt = torch.from_numpy(x).cuda()
    y = torch.zeros_like(xt)
    op.spmv_dev(B, xt, y)
    torch.cuda.synchronize()
    Nt = op.nvars * op.N
    S = ev.S
    Bh = B.cpu().numpy().reshape(g.J, S, Nt, Nt)
    diag, offd = ref.assemble(KIND[form], prm, U)
    scale = np.abs(diag).max()
    assert np.abs(Bh[:, 0] - diag).max() < 1e-12 * scale
    for e in range(g.E):
        ci, co = g.cells[e]
        ti, to = g.trafo[e]
        if co < 0:
            continue
        assert np.abs(Bh[ci, 1 + ti] - offd[e, 0]).max() < 1e-12 * scale
        assert np.abs(Bh[co, 1 + to] - offd[e, 1]).max() < 1e-12 * scale
    # SpMV against the matrix-free operator: linear forms R(x) = A x; Euler J(U) x
    ref_y = ref.evaluate(KIND[form], prm, x) if not form.startswith("euler") else ref.jvp(KIND[form], prm, U, x)
    err = np.abs(y.cpu().numpy() - ref_y).max() / np.abs(ref_y).max()
    say("  colours %d, spmv rel err %.2e" % (ev.ncolors, err))
    assert err < 1e-12
    # y <- alpha A x + beta y (BlockMsrMatrix.SpMV semantics), y non-zero on entry
    y0 = np.random.default_rng(5).uniform(-1, 1, op.ndof)
    y2 = torch.from_numpy(y0).cuda()
    op.spmv_dev(B, xt, y2, alpha=-2.0, beta=0.5)
    want = -2.0 * ref_y + 0.5 * y0
    err2 = np.abs(y2.cpu().numpy() - want).max() / np.abs(want).max()
    say("  alpha/beta spmv rel err %.2e" % err2)
    assert err2 < 1e-12


def test_block_jacobi_and_pcg_on_sip():
    """Block-Jacobi sweep formula and block-Jacobi PCG on the assembled SIP matrix (2D p=2, Dirichlet)."""
    import torch
    from bosss_amd.blocks import BlockJacobi, pcg
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    say("case block jacobi")
    g = CartesianGrid((8, 8))
    op = DifferentialOperator(g, "sip", 2)
    B = op.GetMatrixBuilder().new_matrix()
    op.assemble_dev(torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0"), B, alpha=-1.0)  # -SIP is SPD
    bj = BlockJacobi(op, B, omega=0.7)
    bj.check()
    rng = np.random.default_rng(2)
    b = torch.from_numpy(rng.uniform(-1, 1, op.ndof)).cuda()
    x0 = torch.from_numpy(rng.uniform(-1, 1, op.ndof)).cuda()
    x = x0.clone()
    bj.Solve(x, b, sweeps=1)
    # dense reference of the same sweep
    ref = RefModel(2, (8, 8), 2, 4)
    A = -ref.dense_matrix(OP_SIP, params_vector(1.3, 1.0))
    Nt = op.N
    Dm = np.zeros_like(A)
    for j in range(g.J):
        Dm[j * Nt:(j + 1) * Nt, j * Nt:(j + 1) * Nt] = A[j * Nt:(j + 1) * Nt, j * Nt:(j + 1) * Nt]
    xr = x0.cpu().numpy(); bh = b.cpu().numpy()
    xs = 0.7 * np.linalg.solve(Dm, bh - (A - Dm) @ xr) + 0.3 * xr
    assert np.abs(x.cpu().numpy() - xs).max() < 1e-11 * np.abs(xs).max()
    # PCG with the block-Jacobi preconditioner
    sol, its, rel = pcg(lambda v, out: op.spmv_dev(B, v, out), b, bj.apply_inverse_diagonal, tol=1e-10)
    say("  PCG iterations %d, rel residual %.2e" % (its, rel))
    assert rel <= 1e-10
    assert np.abs(A @ sol.cpu().numpy() - bh).max() < 1e-8 * np.abs(bh).max()


@pytest.mark.parametrize("form,n,p", [("advdiff", 4, 3), ("euler_hllc", 3, 2)])
def test_spmv_tile_autotuner(form, n, p):
    """The apply's autotuner: every candidate panel size gives bitwise the same y (each row sum
    keeps its order), the tuned choice is one of the timed candidates, a shape outside the
    candidates is refused, and the result still matches the oracle."""
    import torch
    from bosss_amd._lib import BosssError
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state
    say("case spmv tiles", form, n, p)
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op = DifferentialOperator(g, form, p, nu=-0.01, velocity=(1.0, 0.5, 0.25))
    ref = RefModel(3, (n,) * 3, p, default_quad_order(form, p), periodic=[1, 1, 1])
    prm = params_vector(1.3, -0.01, (1.0, 0.5, 0.25), 1.4)
    U = euler_state(g, op.N, seed=3) if form.startswith("euler") else np.zeros(op.ndof)
    Ut = torch.from_numpy(U).cuda()
    B = op.GetMatrixBuilder().new_matrix()
    op.assemble_dev(Ut, B)
    x = torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, op.ndof)).cuda()
    y = torch.zeros_like(x)
    op.spmv_dev(B, x, y)                   # first apply: tunes
    info = op.spmv_tiles()
    say("  tiles", info)
    cand = sorted(info["ms"])
    assert len(cand) >= 2 and info["chosen"] in cand
    assert all(info["ms"][c] is not None and info["ms"][c] > 0 for c in cand)
    outs = {}
    for c in cand:
        op.set_spmv_tile(c)
        yc = torch.zeros_like(x)
        op.spmv_dev(B, x, yc)
        outs[c] = yc
    torch.cuda.synchronize()
    for c in cand:
        assert torch.equal(outs[c], y), c
    with pytest.raises(BosssError):
        op.set_spmv_tile(3)
    op.set_spmv_tile(0)                    # tune again at the next apply
    y2 = torch.zeros_like(x)
    op.spmv_dev(B, x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y2, y) and op.spmv_tiles()["chosen"] in cand
    ref_y = ref.evaluate(KIND[form], prm, x.cpu().numpy()) if not form.startswith("euler") else ref.jvp(KIND[form], prm, U, x.cpu().numpy())
    assert np.abs(y.cpu().numpy() - ref_y).max() < 1e-12 * np.abs(ref_y).max()


@pytest.mark.parametrize("shape,per", [((2, 3, 4), True), ((3, 3, 3), False)])
def test_assembly_overwrites_every_slot(shape, per):
    """Slots without a block (a neighbour that an earlier slot already holds, here across a
    periodic direction of two cells) are written too: with beta = 0 storage that held NaN
    becomes a valid matrix, and the apply equals the matrix-free operator."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    say("case assembly into NaN storage", shape, per)
    p, vel = 2, (1.0, 0.5, 0.25)
    g = CartesianGrid(shape, periodic=[per] * 3)
    op = DifferentialOperator(g, "advdiff", p, nu=-0.01, velocity=vel)
    ref = RefModel(3, shape, p, default_quad_order("advdiff", p), periodic=[int(per)] * 3)
    B = torch.full((op.GetMatrixBuilder().entries,), float("nan"), dtype=torch.float64, device="cuda:0")
    op.assemble_dev(torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0"), B)
    x = np.random.default_rng(6).uniform(-1, 1, op.ndof)
    y = torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0")
    op.spmv_dev(B, torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert not torch.isnan(B).any()
    want = ref.evaluate(OP_ADVDIFF, params_vector(1.3, -0.01, vel), x)
    err = np.abs(y.cpu().numpy() - want).max() / np.abs(want).max()
    say("  spmv vs matrix-free: %.2e" % err)
    assert err < 1e-12


def test_spmv_under_graph_capture_before_tuning():
    """An apply first issued inside a captured graph does not time (the autotuner waits for an
    uncaptured call): replays, interleaved with eager work, equal the eager apply bitwise, and
    the later eager call tunes."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    say("case spmv graph capture")
    g = CartesianGrid((4, 4, 4), periodic=[True] * 3)
    op = DifferentialOperator(g, "advdiff", 3, nu=-0.01, velocity=(1.0, 0.5, 0.25))
    B = op.GetMatrixBuilder().new_matrix()
    op.assemble_dev(torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0"), B)
    torch.cuda.synchronize()
    x = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, op.ndof)).cuda()
    yg = torch.zeros_like(x)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            op.spmv_dev(B, x, yg)
    torch.cuda.current_stream().wait_stream(side)
    assert op.spmv_tiles()["chosen"] == 0          # not tuned under capture
    y = torch.zeros_like(x)
    op.spmv_dev(B, x, y)                           # eager: tunes, then applies
    torch.cuda.synchronize()
    assert op.spmv_tiles()["chosen"] != 0
    for rep in range(3):
        yg.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(yg, y), rep
        x.mul_(1.0)                                # eager work between replays
