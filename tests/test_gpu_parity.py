"""GPU parity: the HIP kernels against the CPU oracle on identical seeded inputs.

FP64 bar: 1e-12 relative, measured per field against max|r| of that field and
normwise (per-entry relative error blows up where entries cancel; the reference's own
tables carry ~15 digits).
"""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["tensor", "generic"], autouse=True)
def table_path(request, monkeypatch):
    """Run every case through both kernel families: 1D tensor factors (default) and full LDS tables."""
    if request.param == "generic":
        monkeypatch.setenv("BOSSS_HIP_GENERIC_TABLES", "1")
    else:
        monkeypatch.delenv("BOSSS_HIP_GENERIC_TABLES", raising=False)
    return request.param

KIND = {"sip": OP_SIP, "advdiff": OP_ADVDIFF, "euler_lf": OP_EULER_LF, "euler_hllc": OP_EULER_HLLC}
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def euler_state(J, D, N, detJ, seed):
    rng = np.random.default_rng(seed)
    V = D + 2
    gamma = 1.4
    U = np.zeros((J, V, N))
    rho = rng.uniform(0.8, 1.2, J)
    vel = rng.uniform(-0.1, 0.1, (J, D))
    p = rng.uniform(0.9, 1.1, J)
    E = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)
    s0 = np.sqrt(np.broadcast_to(detJ, (J,))) * 2.0 ** (D / 2.0)     # one detJ, or one per cell
    U[:, 0, 0] = rho * s0
    U[:, 1:1 + D, 0] = rho[:, None] * vel * s0[:, None]
    U[:, D + 1, 0] = E * s0
    U[:, :, 1:] = 0.02 * rng.uniform(-1, 1, (J, V, N - 1)) * s0[:, None, None]
    return U.reshape(-1)


def make_case(form, D, n, p, per, seed=11):
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    order = default_quad_order(form, p)
    nu = -0.01 if form == "advdiff" else 1.0
    vel = (1.0, 0.5, 0.25)[:D]
    g = CartesianGrid((n,) * D, periodic=[per] * D)
    op = DifferentialOperator(g, form, p, penalty=1.3, nu=nu, velocity=vel, gamma=1.4)
    ref = RefModel(D, (n,) * D, p, order, periodic=[int(per)] * D)
    prm = params_vector(1.3, nu, vel, 1.4)
    if form.startswith("euler"):
        U = euler_state(g.J, D, op.N, g.detJ[0], seed)
    else:
        U = np.random.default_rng(seed).uniform(-1, 1, op.ndof)
    return op, ref, prm, U


def rel_err(a, b, V, N):
    a3 = a.reshape(-1, V, N); b3 = b.reshape(-1, V, N)
    per_field = max(np.abs(a3[:, f] - b3[:, f]).max() / max(np.abs(b3[:, f]).max(), 1e-300) for f in range(V))
    normwise = np.linalg.norm(a - b) / np.linalg.norm(b)
    return per_field, normwise


CASES = [
    ("sip", 2, 6, 2, False),      # config-1 kernel (p=2, 3x3 Gauss), Dirichlet boundary
    ("sip", 3, 4, 2, False),      # config-4 bulk kernel
    ("sip", 3, 3, 3, True),
    ("advdiff", 3, 4, 3, True),   # config-2 kernel
    ("advdiff", 3, 3, 2, False),
    ("euler_lf", 3, 4, 2, True),  # config-3 kernels
    ("euler_hllc", 3, 4, 2, True),
    ("euler_lf", 3, 3, 3, True),  # config-5 kernels
    ("euler_hllc", 3, 3, 3, True),
    ("euler_hllc", 2, 5, 2, True),
    ("euler_hllc", 3, 9, 2, True),  # 729 cells = 12 workgroups: XCD block renumbering with a remainder
    ("sip", 2, 30, 2, False),       # 900 cells = 15 workgroups
    ("sip", 2, 5, 1, True),         # the remaining compiled shapes: SIP 2D p = 1 and 3, 3D p = 1,
    ("sip", 2, 4, 3, True),
    ("sip", 3, 3, 1, False),
    ("advdiff", 2, 5, 2, True),     # upwind + SIP 2D p = 2
]


@pytest.mark.parametrize("form,D,n,p,per", CASES)
def test_residual_matches_oracle(form, D, n, p, per, table_path):
    say("case residual", form, D, n, p, per, table_path)
    op, ref, prm, U = make_case(form, D, n, p, per)
    # the family the case asked for is the one that runs (a table check that failed at creation
    # would otherwise send the tensor cases through the generic kernels unnoticed)
    assert op.uses_tensor_tables() == (table_path == "tensor")
    R = op.evaluate(U)
    Rref = ref.evaluate(KIND[form], prm, U)
    pf, nw = rel_err(R, Rref, op.nvars, op.N)
    say("  per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    # alpha/beta accumulation semantics (IEvaluator.Evaluate)
    R0 = np.random.default_rng(3).uniform(-1, 1, op.ndof)
    R2 = op.evaluate(U, alpha=-0.5, beta=2.0, R=R0.copy())
    exp = 2.0 * R0 - 0.5 * Rref
    assert np.abs(R2 - exp).max() < TOL * np.abs(exp).max()


@pytest.mark.parametrize("form,D,n,p,per", CASES)
def test_jvp_vjp_match_oracle(form, D, n, p, per):
    say("case jvp/vjp", form, D, n, p, per)
    op, ref, prm, U = make_case(form, D, n, p, per)
    rng = np.random.default_rng(5)
    dU = rng.uniform(-1, 1, op.ndof) * np.abs(U).max() * 0.1
    W = rng.uniform(-1, 1, op.ndof)
    J = op.jvp(U, dU)
    Jref = ref.jvp(KIND[form], prm, U, dU)
    pf, nw = rel_err(J, Jref, op.nvars, op.N)
    say("  jvp per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL
    JT = op.vjp(U, W)
    JTref = ref.vjp(KIND[form], prm, U, W)
    pf, nw = rel_err(JT, JTref, op.nvars, op.N)
    say("  vjp per-field %.2e normwise %.2e" % (pf, nw))
    assert pf < TOL and nw < TOL


@pytest.mark.parametrize("form,D,n,p,per", [CASES[0], CASES[3], CASES[8]])
def test_device_entry_points_and_fused_residual_vjp(form, D, n, p, per):
    import torch
    say("case device", form, D, n, p, per)
    op, ref, prm, U = make_case(form, D, n, p, per)
    W = np.random.default_rng(9).uniform(-1, 1, op.ndof)
    dev = torch.device("cuda:0")
    Ut = torch.from_numpy(U).to(dev)
    Wt = torch.from_numpy(W).to(dev)
    R1 = torch.zeros_like(Ut); G1 = torch.zeros_like(Ut)
    R2 = torch.full_like(Ut, 7.0); G2 = torch.zeros_like(Ut)
    J1 = torch.zeros_like(Ut)
    op.evaluate_dev(Ut, R1)
    op.vjp_dev(Ut, Wt, G1)
    op.jvp_dev(Ut, Wt, J1)
    # consume the results with torch ops on the same (current) stream, no explicit sync:
    # the kernels must be ordered on torch's stream
    S1 = (R1 * 2.0).cpu().numpy() / 2.0
    Jref = ref.jvp(KIND[form], prm, U, W)
    pf, nw = rel_err(J1.cpu().numpy(), Jref, op.nvars, op.N)
    assert pf < TOL and nw < TOL
    op.evaluate_vjp_dev(Ut, R2, Wt, G2, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    Rref = ref.evaluate(KIND[form], prm, U)
    Gref = ref.vjp(KIND[form], prm, U, W)
    pf, nw = rel_err(S1, Rref, op.nvars, op.N)
    assert pf < TOL and nw < TOL
    for a, b in ((R1, Rref), (R2, Rref), (G1, Gref), (G2, Gref)):
        pf, nw = rel_err(a.cpu().numpy(), b, op.nvars, op.N)
        assert pf < TOL and nw < TOL
    # each kernel is deterministic: no atomics, fixed summation order
    R3 = torch.zeros_like(Ut); G3 = torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R3, Wt, G3, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    assert torch.equal(R2, R3) and torch.equal(G2, G3)


def graded_nodes(D, n, seed=5):
    """Increasing nodes on [0, 1]: n + d cells in direction d, widths varying by up to 3x."""
    rng = np.random.default_rng(seed)
    out = []
    for d in range(D):
        w = rng.uniform(1.0, 3.0, n + d)
        out.append(np.concatenate([[0.0], np.cumsum(w) / w.sum()]))
    return out


GRADED = [
    ("sip", 2, 5, 2, False),
    ("sip", 3, 3, 2, False),
    ("advdiff", 3, 3, 3, True),
    ("euler_lf", 3, 3, 2, True),
    ("euler_hllc", 3, 3, 2, True),
    ("euler_hllc", 3, 3, 3, True),
    ("euler_hllc", 2, 4, 2, True),
]


@pytest.mark.parametrize("form,D,n,p,per", GRADED)
def test_graded_grid_matches_oracle(form, D, n, p, per):
    """Graded tensor grids (Cartesian3DGrid over node arrays): every cell has its own widths
    and the cell counts differ per direction, so per-cell metrics, the scaling of each side
    of a face and the direction of every inverse-Jacobian entry are exercised.  Residual,
    JVP, VJP and the fused device residual + VJP against the oracle on the same grid."""
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator, default_quad_order
    say("case graded", form, D, n, p, per)
    nodes = graded_nodes(D, n)
    nu = -0.01 if form == "advdiff" else 1.0
    vel = (1.0, 0.5, 0.25)[:D]
    g = CartesianGrid(nodes=nodes, periodic=[per] * D)
    op = DifferentialOperator(g, form, p, penalty=1.3, nu=nu, velocity=vel, gamma=1.4)
    ref = RefModel(D, None, p, default_quad_order(form, p), periodic=[int(per)] * D, nodes=nodes)
    assert ref.J == g.J and ref.E == len(g.cells)
    prm = params_vector(1.3, nu, vel, 1.4)
    rng = np.random.default_rng(13)
    U = euler_state(g.J, D, op.N, g.detJ, 17) if form.startswith("euler") else rng.uniform(-1, 1, op.ndof)
    dU = rng.uniform(-1, 1, op.ndof) * np.abs(U).max() * 0.1
    W = rng.uniform(-1, 1, op.ndof)
    K = KIND[form]
    Rref, Jref, Gref = ref.evaluate(K, prm, U), ref.jvp(K, prm, U, dU), ref.vjp(K, prm, U, W)
    for name, got, want in (("residual", op.evaluate(U), Rref), ("jvp", op.jvp(U, dU), Jref), ("vjp", op.vjp(U, W), Gref)):
        pf, nw = rel_err(got, want, op.nvars, op.N)
        say("  %s per-field %.2e normwise %.2e" % (name, pf, nw))
        assert pf < TOL and nw < TOL
    dev = torch.device("cuda:0")
    Ut, Wt = torch.from_numpy(U).to(dev), torch.from_numpy(W).to(dev)
    R, G = torch.zeros_like(Ut), torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R, Wt, G)
    torch.cuda.synchronize()
    for name, got, want in (("fused residual", R, Rref), ("fused vjp", G, Gref)):
        pf, nw = rel_err(got.cpu().numpy(), want, op.nvars, op.N)
        say("  %s per-field %.2e normwise %.2e" % (name, pf, nw))
        assert pf < TOL and nw < TOL


def test_check_finite_reports_the_first_nan():
    """CheckForNanOrInf (DifferentialOperator.cs:1602,1623,1658): off by default; on, a NaN injected
    into one coefficient of U makes the residual, JVP, VJP and fused calls raise BOSSS_ERR_NONFINITE
    naming a non-finite output entry, and a finite state passes."""
    import torch
    from bosss_amd._lib import BosssError
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    from bosss_amd.workloads import euler_state
    print("case check_finite euler_hllc p=2 3^3", flush=True)
    g = CartesianGrid((3, 3, 3), periodic=[True] * 3)
    op = DifferentialOperator(g, "euler_hllc", 2)
    U = torch.from_numpy(euler_state(g, op.N, seed=1)).cuda()
    W = torch.rand_like(U)
    R, G = torch.zeros_like(U), torch.zeros_like(U)
    Ubad = U.clone()
    Ubad[7 * op.nvars * op.N + 3] = float("nan")     # cell 7, variable 0, mode 3
    op.evaluate_dev(Ubad, R)                          # unchecked: no error
    torch.cuda.synchronize()
    op.set_check_finite(True)
    op.evaluate_dev(U, R)
    op.evaluate_vjp_dev(U, R, W, G)
    for call in (lambda: op.evaluate_dev(Ubad, R), lambda: op.jvp_dev(Ubad, W, G), lambda: op.vjp_dev(Ubad, W, G),
                 lambda: op.evaluate_vjp_dev(Ubad, R, W, G)):
        with pytest.raises(BosssError) as e:
            call()
        assert e.value.code == -4 and "NaN or Inf" in str(e.value)
    with pytest.raises(BosssError):
        op.evaluate(Ubad.cpu().numpy())
    op.set_check_finite(False)
    op.close()


def euler_state_supersonic(J, D, N, detJ, seed):
    """Cell states from rest to about Mach 3 in random directions, density and pressure varying
    up to 6.7x between neighbours, higher modes at 0.2 % of each variable's scale (so the
    pressure stays positive at every node): the faces see all four HLLC regions, shocks and
    rarefactions on either side, and PVRS estimates clipped at zero.  Returns the coefficients
    and the cell means (rho, m, E)."""
    rng = np.random.default_rng(seed)
    V = D + 2
    gamma = 1.4
    rho = rng.uniform(0.3, 2.0, J)
    p = rng.uniform(0.3, 2.0, J)
    a = np.sqrt(gamma * p / rho)
    vel = rng.uniform(-3.0, 3.0, (J, D)) * (a / np.sqrt(D))[:, None]
    mean = np.zeros((J, V))
    mean[:, 0] = rho
    mean[:, 1:1 + D] = rho[:, None] * vel
    mean[:, D + 1] = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)
    scale = np.concatenate([rho[:, None], np.repeat((rho * a)[:, None], D, 1), mean[:, D + 1:]], 1)
    s0 = np.sqrt(np.broadcast_to(detJ, (J,))) * 2.0 ** (D / 2.0)
    U = np.zeros((J, V, N))
    U[:, :, 0] = mean * s0[:, None]
    U[:, :, 1:] = 0.002 * rng.uniform(-1, 1, (J, V, N - 1)) * (scale * s0[:, None])[:, :, None]
    return U.reshape(-1), mean


@pytest.mark.parametrize("form", ["euler_hllc", "euler_lf"])
@pytest.mark.parametrize("D,n,p", [(3, 4, 2), (3, 4, 3), (2, 6, 2)])
def test_euler_every_flux_region_matches_oracle(form, D, n, p):
    """Euler fluxes on supersonic, strongly discontinuous states.  The random states of the other
    cases are at Mach 0.1 with 2 % jumps, so only the subsonic star regions of HLLC run there;
    here the host classifies every face by the textbook HLLC of the cell means (Toro 10.4-10.6)
    and the case must contain each region: F_L (S_L >= 0), F*_L, F*_R and F_R (S_R <= 0), with
    shocks.  Residual, JVP, VJP and the fused device residual + VJP against the oracle."""
    import torch
    from test_oracle_jumps import toro_hllc
    say("case supersonic", form, D, n, p)
    op, ref, prm, _ = make_case(form, D, n, p, True)
    g = op.grid
    U, mean = euler_state_supersonic(g.J, D, op.N, g.detJ[0], 23)
    regions = {"L": 0, "*L": 0, "*R": 0, "R": 0}
    shocks = vacuum = 0
    for (ci, co), nrm in zip(g.cells, g.normals):
        _, info = toro_hllc(mean[ci], mean[co], nrm)
        regions[info["branch"]] += 1
        shocks += info["shockL"] or info["shockR"]
        vacuum += info["pstar"] == 0.0
    say("  regions %s, shocks %d, clipped PVRS %d of %d faces" % (regions, shocks, vacuum, g.E))
    assert min(regions.values()) >= 3 and shocks >= 3
    rng = np.random.default_rng(29)
    dU = rng.uniform(-1, 1, op.ndof) * np.abs(U).max() * 0.1
    W = rng.uniform(-1, 1, op.ndof)
    K = KIND[form]
    Rref, Jref, Gref = ref.evaluate(K, prm, U), ref.jvp(K, prm, U, dU), ref.vjp(K, prm, U, W)
    assert np.isfinite(Rref).all() and np.isfinite(Jref).all() and np.isfinite(Gref).all()
    for name, got, want in (("residual", op.evaluate(U), Rref), ("jvp", op.jvp(U, dU), Jref), ("vjp", op.vjp(U, W), Gref)):
        pf, nw = rel_err(got, want, op.nvars, op.N)
        say("  %s per-field %.2e normwise %.2e" % (name, pf, nw))
        assert pf < TOL and nw < TOL
    dev = torch.device("cuda:0")
    Ut, Wt = torch.from_numpy(U).to(dev), torch.from_numpy(W).to(dev)
    R, G = torch.zeros_like(Ut), torch.zeros_like(Ut)
    op.evaluate_vjp_dev(Ut, R, Wt, G)
    torch.cuda.synchronize()
    for name, got, want in (("fused residual", R, Rref), ("fused vjp", G, Gref)):
        pf, nw = rel_err(got.cpu().numpy(), want, op.nvars, op.N)
        say("  %s per-field %.2e normwise %.2e" % (name, pf, nw))
        assert pf < TOL and nw < TOL
