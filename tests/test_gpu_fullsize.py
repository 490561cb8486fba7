"""The measured workloads at full size: Euler HLLC p = 3 on 64^3 (26 214 400 DOF), upwind
advection + SIP p = 3 on 64^3, XDG Poisson on the 32^3 sphere, SIP p = 2 on 64^2 and Euler
LF / HLLC p = 2 on 32^3.

The oracle finishes this size in a few seconds on the host's cores, so the fused residual +
VJP kernel and the FP64 JVP kernel are compared with it entry by entry (1e-12 per field, as
the small parity cases).  Size-independent properties are checked beside it: conservation
of the mean modes on the periodic grid, free-stream preservation, the adjoint identity
between the JVP and VJP kernels, the fused kernel against the separate residual and VJP
kernels, and bitwise run-to-run determinism.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
N_CELLS = 64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b, V, N):
    a3 = a.reshape(-1, V, N); b3 = b.reshape(-1, V, N)
    per_field = max(np.abs(a3[:, f] - b3[:, f]).max() / max(np.abs(b3[:, f]).max(), 1e-300) for f in range(V))
    return per_field, np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def case():
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state, uniform_coefficients
    from oracle.bosss_ref import RefModel, params_vector
    n, p = N_CELLS, 3
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op = DifferentialOperator(g, "euler_hllc", p, gamma=1.4)
    assert op.uses_tensor_tables()   # the headline kernels, not the generic-table fallback
    ref = RefModel(3, (n,) * 3, p, default_quad_order("euler_hllc", p), periodic=[1, 1, 1])
    prm = params_vector(1.3, 1.0, (1.0, 0.5, 0.25), 1.4)
    U = euler_state(g, op.N, seed=0)
    W = uniform_coefficients(op.ndof, 1)
    dev = torch.device("cuda:0")
    return dict(g=g, op=op, ref=ref, prm=prm, U=U, W=W, Ut=torch.from_numpy(U).to(dev), Wt=torch.from_numpy(W).to(dev))


def test_fused_residual_vjp_against_oracle_at_full_size(case):
    import torch
    from oracle.bosss_ref import OP_EULER_HLLC
    op, ref, prm, U, W, Ut, Wt = (case[k] for k in ("op", "ref", "prm", "U", "W", "Ut", "Wt"))
    V, N = op.nvars, op.N
    say("case fullsize fused euler_hllc p=3 64^3")
    R = torch.zeros_like(Ut); G = torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R, Wt, G, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    Rh, Gh = R.cpu().numpy(), G.cpu().numpy()
    say("  oracle residual + VJP")
    Rref = ref.evaluate(OP_EULER_HLLC, prm, U)
    Gref = ref.vjp(OP_EULER_HLLC, prm, U, W)
    pf, nw = rel_err(Rh, Rref, V, N)
    say("  residual per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    pf, nw = rel_err(Gh, Gref, V, N)
    say("  VJP      per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    # the separate residual and VJP kernels give the same result
    R1 = torch.zeros_like(Ut); G1 = torch.zeros_like(Ut)
    op.evaluate_dev(Ut, R1)
    op.vjp_dev(Ut, Wt, G1)
    torch.cuda.synchronize()
    for a, b in ((R1, Rh), (G1, Gh)):
        pf, nw = rel_err(a.cpu().numpy(), b, V, N)
        assert pf < TOL and nw < TOL
    # bitwise deterministic: no atomics, fixed summation order
    R2 = torch.zeros_like(Ut); G2 = torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R2, Wt, G2, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    assert torch.equal(R, R2) and torch.equal(G, G2)
    # conservation: on the periodic grid every face flux enters two mean-mode residuals with
    # opposite signs and the volume term of the constant test function vanishes
    R3 = Rh.reshape(-1, V, N)
    for v in range(V):
        s, scale = R3[:, v, 0].sum(), np.abs(R3[:, v, 0]).sum()
        say("  conservation field %d: |sum| / sum|.| = %.2e" % (v, abs(s) / scale))
        assert abs(s) <= 1e-12 * scale


def test_jvp_against_oracle_and_adjoint_identity_at_full_size(case):
    import torch
    from oracle.bosss_ref import OP_EULER_HLLC
    op, ref, prm, U, W, Ut, Wt = (case[k] for k in ("op", "ref", "prm", "U", "W", "Ut", "Wt"))
    V, N = op.nvars, op.N
    say("case fullsize jvp euler_hllc p=3 64^3")
    dU = np.random.default_rng(5).uniform(-1, 1, op.ndof)
    dUt = torch.from_numpy(dU).cuda()
    JdU = torch.zeros_like(Ut); JtW = torch.zeros_like(Ut)
    op.jvp_dev(Ut, dUt, JdU)
    op.vjp_dev(Ut, Wt, JtW)
    torch.cuda.synchronize()
    say("  oracle JVP")
    Jref = ref.jvp(OP_EULER_HLLC, prm, U, dU)
    pf, nw = rel_err(JdU.cpu().numpy(), Jref, V, N)
    say("  JVP per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    # <W, J dU> = <J^T W, dU>: the forward-mode JVP kernel against the hand-written adjoint
    lhs = float(torch.dot(Wt, JdU))
    rhs = float(torch.dot(JtW, dUt))
    scale = float(torch.dot(Wt.abs(), JdU.abs()))
    say("  adjoint identity: %.15e vs %.15e (rel %.2e)" % (lhs, rhs, abs(lhs - rhs) / scale))
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_free_stream_is_preserved_at_full_size(case):
    import torch
    g, op = case["g"], case["op"]
    V, N = op.nvars, op.N
    say("case fullsize free stream")
    # a constant state: only the mean mode is set (normalised mode 0 = 1 / sqrt(cell volume))
    rho, vel, p, gamma = 1.1, np.array([0.07, -0.03, 0.05]), 0.95, 1.4
    state = np.concatenate([[rho], rho * vel, [p / (gamma - 1) + 0.5 * rho * vel @ vel]])
    Uc = np.zeros((g.J, V, N))
    Uc[:, :, 0] = state[None, :] * np.sqrt(g.volume)[:, None]
    Uct = torch.from_numpy(Uc.reshape(-1)).cuda()
    Rc = torch.zeros_like(Uct)
    op.evaluate_dev(Uct, Rc)
    R = torch.zeros_like(Uct)
    op.evaluate_dev(case["Ut"], R)
    torch.cuda.synchronize()
    worst = float(Rc.abs().max()) / float(R.abs().max())
    say("  max |R(constant)| / max |R(U)| = %.2e" % worst)
    assert worst <= 1e-11


def test_advdiff_p3_64cube_fp64_fp32_and_assembled_at_full_size():
    """Upwind advection + SIP (nu = 0.01), 64^3 periodic, p = 3 (5.2 M DOF): the matrix-free
    residual in FP64 against the oracle (1e-12) and in FP32 (error stated, bound 1e-5), and
    the assembled block-ELL apply M U + b against the matrix-free residual."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import uniform_coefficients
    from oracle.bosss_ref import RefModel, params_vector, OP_ADVDIFF
    say("case fullsize advdiff p=3 64^3")
    n, p, nu, vel = N_CELLS, 3, -0.01, (1.0, 0.5, 0.25)
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op = DifferentialOperator(g, "advdiff", p, penalty=1.3, nu=nu, velocity=vel)
    op32 = DifferentialOperator(g, "advdiff", p, penalty=1.3, nu=nu, velocity=vel, precision="fp32")
    U = uniform_coefficients(op.ndof, 3)
    Ut = torch.from_numpy(U).cuda()
    R = torch.zeros_like(Ut)
    op.evaluate_dev(Ut, R)
    R32 = torch.zeros(op.ndof, dtype=torch.float32, device=Ut.device)
    op32.evaluate_dev(Ut.float(), R32)
    torch.cuda.synchronize()
    ref = RefModel(3, (n,) * 3, p, default_quad_order("advdiff", p), periodic=[1, 1, 1])
    Rref = ref.evaluate(OP_ADVDIFF, params_vector(1.3, nu, vel, 1.4), U)
    pf, nw = rel_err(R.cpu().numpy(), Rref, 1, op.N)
    say("  FP64 per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    pf32, nw32 = rel_err(R32.double().cpu().numpy(), Rref, 1, op.N)
    say("  FP32 per-field %.2e normwise %.2e" % (pf32, nw32))
    assert pf32 < 1e-5
    # the operator is linear: R(U) = M U + b with the assembled matrix (b = R(0))
    ev = op.GetMatrixBuilder()
    B = ev.new_matrix()
    op.assemble_dev(torch.zeros_like(Ut), B)
    y = torch.zeros_like(Ut)
    op.evaluate_dev(torch.zeros_like(Ut), y)
    op.spmv_dev(B, Ut, y, alpha=1.0, beta=1.0)
    torch.cuda.synchronize()
    d = float((y - R).abs().max()) / float(R.abs().max())
    say("  |M U + b - R(U)| / |R| = %.2e" % d)
    assert d < 1e-12
    del B


def test_xdg_sphere_32cube_assembly_and_solves_at_full_size():
    """XDG Poisson on the 32^3 sphere (p = 2, 345 k DOF): GPU assembly against the oracle's
    restated assembly (1e-12), and both block-Jacobi PCG solvers to ||b - M x|| <= 1e-10 ||b||
    with the residual recomputed from the oracle's matrix."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.xdg import QuadLevelSet, XdgRules, XdgOperator
    from oracle.bosss_ref import xdg_assemble, xdg_to_csr
    say("case fullsize xdg sphere 32^3")
    g = CartesianGrid((32,) * 3)
    rules = XdgRules(g, QuadLevelSet.sphere([0.5] * 3, 0.3), n=3)
    mu, src = (-1.0, -10.0), (1.0, 1.0)
    op = XdgOperator(rules, 2, mu, source=src, penalty=2.0, g_D=None)
    B = op.new_matrix()
    a = torch.zeros(op.ndof, dtype=torch.float64, device="cuda:0")
    op.assemble_dev(B, a)
    torch.cuda.synchronize()
    Bref, aref, lay = xdg_assemble(rules, 2, mu, src, penalty=2.0, gD=rules.boundary_values(None))
    eb = np.abs(B.cpu().numpy() - Bref).max() / np.abs(Bref).max()
    ea = np.abs(a.cpu().numpy() - aref).max() / np.abs(aref).max()
    say("  blocks %.2e affine %.2e" % (eb, ea))
    assert eb < 1e-12 and ea < 1e-12
    M = xdg_to_csr(rules, lay, Bref)
    b = -a
    bh = b.cpu().numpy()
    for prec in ("fp64", "fp32"):
        x = torch.zeros_like(a)
        it, outer, rel = op.solve_dev(B, b, x, tol=1e-10, maxit=50000, precision=prec)
        torch.cuda.synchronize()
        true_rel = np.linalg.norm(bh - M @ x.cpu().numpy()) / np.linalg.norm(bh)
        say("  %s: %d iterations, %d outer, reported %.2e, oracle-matrix residual %.2e" % (prec, it, outer, rel, true_rel))
        assert rel <= 1e-10 and true_rel <= 2e-10
    op.close()


@pytest.mark.parametrize("form,D,n,p,per", [("sip", 2, 64, 2, False), ("euler_lf", 3, 32, 2, True), ("euler_hllc", 3, 32, 2, True)])
def test_small_configurations_at_full_size(form, D, n, p, per):
    """SIP Poisson on 64^2 (p = 2, Dirichlet) and Euler LF / HLLC on 32^3 (p = 2, periodic):
    the fused residual + VJP kernel against the oracle (1e-12 per field)."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    from bosss_amd.workloads import euler_state, uniform_coefficients
    from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_EULER_LF, OP_EULER_HLLC
    say("case fullsize", form, D, n, p, per)
    kind = {"sip": OP_SIP, "euler_lf": OP_EULER_LF, "euler_hllc": OP_EULER_HLLC}[form]
    g = CartesianGrid((n,) * D, periodic=[per] * D)
    op = DifferentialOperator(g, form, p, penalty=1.3, nu=1.0, gamma=1.4)
    ref = RefModel(D, (n,) * D, p, default_quad_order(form, p), periodic=[int(per)] * D)
    prm = params_vector(1.3, 1.0, (1.0, 0.5, 0.25)[:D], 1.4)
    U = euler_state(g, op.N, seed=4) if form.startswith("euler") else uniform_coefficients(op.ndof, 4)
    W = uniform_coefficients(op.ndof, 6)
    Ut, Wt = torch.from_numpy(U).cuda(), torch.from_numpy(W).cuda()
    R = torch.zeros_like(Ut); G = torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R, Wt, G, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    for name, got, want in (("residual", R, ref.evaluate(kind, prm, U)), ("VJP", G, ref.vjp(kind, prm, U, W))):
        pf, nw = rel_err(got.cpu().numpy(), want, op.nvars, op.N)
        say("  %s per-field %.2e normwise %.2e" % (name, pf, nw))
        assert pf < TOL and nw < TOL
