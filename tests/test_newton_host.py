"""Host logic of the Newton-Krylov driver and GMRES, on the CPU.

The operator is a stand-in that evaluates the residual and its JVP/VJP with the
CPU oracle on CPU tensors, so the driver (GMRES, defect correction, adjoint solve)
is checked without a GPU.  The GPU runs of the same driver are in test_gpu_newton.py.
Also the host-side partitioning of the Schwarz smoother (its GPU runs: test_gpu_schwarz.py).
"""
import numpy as np
import torch

from bosss_amd.grid import CartesianGrid
from bosss_amd.newton import ImplicitStep, gmres
from bosss_amd.workloads import euler_state
from oracle.bosss_ref import RefModel, params_vector, OP_EULER_HLLC


class OracleOp:
    def __init__(self, n, p):
        self.ref = RefModel(3, (n,) * 3, p, 3 * p, periodic=[1, 1, 1])
        self.prm = params_vector(gamma=1.4)

    def _run(self, fn, *args, out):
        res = fn(OP_EULER_HLLC, self.prm, *[a.double().numpy() for a in args])
        out.copy_(torch.from_numpy(res).to(out.dtype))

    def evaluate_dev(self, U, R):
        self._run(self.ref.evaluate, U, out=R)

    def jvp_dev(self, U, dU, out):
        self._run(self.ref.jvp, U, dU, out=out)

    def vjp_dev(self, U, W, out):
        self._run(self.ref.vjp, U, W, out=out)


def test_gmres_solves_small_nonsymmetric_system():
    rng = np.random.default_rng(0)
    A = torch.from_numpy(np.eye(40) * 4 + rng.uniform(-1, 1, (40, 40)))
    b = torch.from_numpy(rng.uniform(-1, 1, 40))
    x, its, rel = gmres(lambda v, out: out.copy_(A @ v), b, 1e-12, restart=15, maxit=200)
    assert rel <= 1e-12 and torch.linalg.vector_norm(A @ x - b) <= 1e-11 * torch.linalg.vector_norm(b)


def test_gmres_estimate_mode_saves_one_apply_and_returns_the_same_solution():
    """true_residual=False (Newton's inner solves): the converged cycle's least-squares
    estimate is returned instead of an extra apply; the solution is the same."""
    rng = np.random.default_rng(1)
    A = torch.from_numpy(np.eye(60) * 5 + rng.uniform(-1, 1, (60, 60)))
    b = torch.from_numpy(rng.uniform(-1, 1, 60))
    calls = [0]

    def apply(v, out):
        calls[0] += 1
        out.copy_(A @ v)

    x1, it1, rel1 = gmres(apply, b, 1e-8, restart=50, maxit=200)
    n1, calls[0] = calls[0], 0
    x2, it2, rel2 = gmres(apply, b, 1e-8, restart=50, maxit=200, true_residual=False)
    n2 = calls[0]
    assert it1 == it2 and n2 == n1 - 1 and torch.equal(x1, x2)
    true = float(torch.linalg.vector_norm(A @ x2 - b) / torch.linalg.vector_norm(b))
    assert rel2 <= 1e-8 and abs(rel2 - true) <= 1e-3 * true + 1e-15
    # a cycle that ends unconverged still restarts from the true residual
    x3, it3, rel3 = gmres(apply, b, 1e-8, restart=5, maxit=400, true_residual=False)
    assert float(torch.linalg.vector_norm(A @ x3 - b) / torch.linalg.vector_norm(b)) <= 1.01e-8


def test_newton_fp64_and_defect_correction_converge():
    n, p = 4, 2
    g = CartesianGrid((n,) * 3, periodic=[True] * 3)
    Un = torch.from_numpy(euler_state(g, 10, seed=7))
    dt = 0.5 / n / (2 * p + 1)
    st = ImplicitStep(OracleOp(n, p), Un, dt, op32=OracleOp(n, p))
    U64 = st.solve_fp64(Un, tol=1e-10)
    assert st.stats["history"][-1] <= 1e-10
    Um = st.solve_mixed(Un, tol=1e-10)
    assert st.stats["history"][-1] <= 1e-10
    assert torch.linalg.vector_norm(U64 - Um) <= 1e-9


class AtanOp:
    """Stand-in operator R(U) = atan(U) elementwise: Newton's method on atan diverges from
    |x0| > 1.39 with full steps and converges with a line search."""

    def evaluate_dev(self, U, R):
        torch.atan(U, out=R)

    def jvp_dev(self, U, dU, out):
        out.copy_(dU / (1.0 + U * U))


def test_parab3p_minimises_the_parabola_and_is_safeguarded():
    from bosss_amd.newton import parab3p
    ff = lambda lam: (lam - 0.3) ** 2 + 1.0
    assert abs(parab3p(1.0, 2.0, ff(0.0), ff(1.0), ff(2.0)) - 0.3) < 1e-14      # inside [0.1, 0.5]
    assert parab3p(0.5, 1.0, ff(0.0), ff(0.5), ff(1.0)) == 0.25                    # clipped to sigma1 * lambdac
    ff2 = lambda lam: (lam - 0.01) ** 2
    assert parab3p(1.0, 2.0, ff2(0.0), ff2(1.0), ff2(2.0)) == 0.1                  # clipped to sigma0 * lambdac
    concave = lambda lam: -lam * lam
    assert parab3p(1.0, 2.0, concave(0.0), concave(1.0), concave(2.0)) == 0.5      # no minimum: sigma1 * lambdac


def test_line_search_damps_a_diverging_newton_iteration():
    """F(U) = atan(U) (dt -> infinity): full Newton steps from U0 = 10 diverge and the driver
    reports converged = False; the Armijo line search (Newton.cs:807-860) converges to 0."""
    Un = torch.full((8,), 10.0, dtype=torch.float64)
    st = ImplicitStep(AtanOp(), Un, 1e30)
    U = st.solve_fp64(Un, tol=1e-12, max_newton=12, globalization="none")
    assert not st.stats["converged"] and st.stats["newton"] == 12
    assert not np.isfinite(st.stats["history"][-1]) or st.stats["history"][-1] > st.stats["history"][0]
    U = st.solve_fp64(Un, tol=1e-12, max_newton=40)
    s = st.stats
    assert s["converged"] and float(U.abs().max()) < 1e-12
    assert s["step_lengths"][0] < 1.0                    # the first step was damped
    hist = s["history"]
    assert all(hist[i + 1] < (1 - 1e-4 * s["step_lengths"][i]) * hist[i] for i in range(len(hist) - 1))


def test_cartesian_parts_cover_every_cell_once():
    """Schwarz parts of a Cartesian grid: boxes of agg cells per direction (the last box of a
    direction takes the remainder), every cell in exactly one part, each part a box."""
    from bosss_amd.blocks import cartesian_parts
    g = CartesianGrid((5, 4, 3))
    parts = cartesian_parts(g, (2, 2, 2))
    assert len(parts) == 3 * 2 * 2
    allc = np.sort(np.concatenate(parts))
    assert (allc == np.arange(g.J)).all()
    idx = np.indices(g.ncells).reshape(3, -1).T
    sizes = sorted(len(p) for p in parts)
    assert sizes == sorted([2 * 2 * 2] * 4 + [2 * 2 * 1] * 4 + [1 * 2 * 2] * 2 + [1 * 2 * 1] * 2)
    for p in parts:
        box = idx[p]
        ext = box.max(0) - box.min(0) + 1
        assert np.prod(ext) == len(p)                # a full box of cells
