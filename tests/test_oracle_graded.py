"""Pins of the oracle on graded tensor grids (Cartesian2DGrid/Cartesian3DGrid over explicit
node arrays, Grid3D.cs:85-235): every cell has its own widths, so per-cell metrics, the
scaling of each side of a face and the direction of each inverse-Jacobian entry matter.

Checks: metrics in closed form (detJ, invJ, CellLengthScale GridData_CellData.cs:681-684,
sqrt(Gram) of each face, the sealing sum of each cell, DerivativeTestMain.cs:1277-1333); the
uniform node path equals the box constructor; exact Laplacian / advection-diffusion of
quadratics in interior cells (DerivativeTestMain.cs:694-728); zero residual of a constant
Euler state and conservation of the cell-mean sums on a periodic graded grid.
"""
import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC


def graded_nodes(D, n, seed=5, periodic=False):
    """Increasing nodes on [0, 1] per direction with widths varying by up to 3x."""
    rng = np.random.default_rng(seed)
    out = []
    for d in range(D):
        w = rng.uniform(1.0, 3.0, n + d)          # a different cell count per direction
        out.append(np.concatenate([[0.0], np.cumsum(w) / w.sum()]))
    return out


def cell_boxes(nodes):
    D = len(nodes)
    shape = tuple(len(x) - 1 for x in nodes)
    idx = np.indices(shape).reshape(D, -1).T       # x slowest, last index fastest
    lo = np.stack([nodes[d][idx[:, d]] for d in range(D)], 1)
    hi = np.stack([nodes[d][idx[:, d] + 1] for d in range(D)], 1)
    return idx, lo, hi


def project(m, nodes, funcs):
    """L2-projection of point functions onto the orthonormal DG basis of each cell."""
    g, t = m.grid(), m.tables()
    _, lo, hi = cell_boxes(nodes)
    U = np.zeros((m.J, len(funcs), m.N))
    for j in range(m.J):
        x = 0.5 * (lo[j] + hi[j]) + 0.5 * (hi[j] - lo[j]) * t["vnodes"]
        for f, fn in enumerate(funcs):
            U[j, f] = (t["vw"] * fn(x) * g["detJ"][j] * g["scaling"][j]) @ t["phi"]
    return U.reshape(-1)


def interior(nodes, layers=1):
    idx, _, _ = cell_boxes(nodes)
    n = np.array([len(x) - 1 for x in nodes])
    return np.all((idx >= layers) & (idx < n - layers), axis=1)


@pytest.mark.parametrize("D,periodic", [(2, False), (3, False), (3, True)])
def test_graded_metrics_closed_form(D, periodic):
    nodes = graded_nodes(D, 4)
    m = RefModel(D, None, 2, 4, periodic=[int(periodic)] * D, nodes=nodes)
    g = m.grid()
    idx, lo, hi = cell_boxes(nodes)
    h = hi - lo
    assert m.J == len(h)
    assert np.abs(g["detJ"] - np.prod(h / 2, 1)).max() < 1e-16
    assert np.abs(g["scaling"] - 1 / np.sqrt(np.prod(h / 2, 1))).max() < 1e-12 * g["scaling"].max()
    for d in range(D):
        assert np.abs(g["invJ"][:, d, d] - 2 / h[:, d]).max() < 1e-12 * (2 / h).max()
    off = g["invJ"].copy()
    for d in range(D):
        off[:, d, d] = 0
    assert not off.any()
    # CellLengthScale = 2 |K| / (surface + boundary surface), boundary faces counted twice
    farea = np.stack([np.prod(np.delete(h, d, 1), 1) for d in range(D)], 1)
    n = np.array([len(x) - 1 for x in nodes])
    bnd = np.zeros(m.J)
    if not periodic:
        for d in range(D):
            bnd += farea[:, d] * ((idx[:, d] == 0).astype(float) + (idx[:, d] == n[d] - 1))
    ls = 2 * np.prod(h, 1) / (2 * farea.sum(1) + bnd)
    assert np.abs(g["lenscale"] - ls).max() < 1e-14
    # each face: sqrt(Gram) = area / 2^(D-1) of both neighbours; sealing sum per cell = 0
    seal = np.zeros((m.J, D))
    for e in range(m.E):
        ci, co = g["cells"][e]
        d = int(np.argmax(np.abs(g["normals"][e])))
        a = g["sqrtgram"][e] * 2 ** (D - 1)
        assert abs(a - farea[ci, d]) < 1e-15
        seal[ci] += a * g["normals"][e]
        if co >= 0:
            assert abs(a - farea[co, d]) < 1e-15
            seal[co] -= a * g["normals"][e]
    assert np.abs(seal).max() < 1e-15


@pytest.mark.parametrize("kind", [OP_SIP, OP_ADVDIFF, OP_EULER_HLLC])
def test_uniform_nodes_match_the_box_constructor(kind):
    D, n, p = 3, 3, 2
    m0 = RefModel(D, (n,) * D, p, 6, periodic=[1] * D)
    m1 = RefModel(D, None, p, 6, periodic=[1] * D, nodes=[np.arange(n + 1) / n] * D)
    V = m0.nvars(kind)
    U = np.random.default_rng(2).uniform(0.5, 1.0, m0.J * V * m0.N)
    prm = params_vector(1.3, 0.1, (1.0, 0.5, 0.25), 1.4)
    r0, r1 = m0.evaluate(kind, prm, U), m1.evaluate(kind, prm, U)
    assert np.abs(r0 - r1).max() < 1e-13 * np.abs(r0).max()


@pytest.mark.parametrize("D,p", [(2, 2), (3, 2), (3, 3)])
def test_graded_sip_laplacian_of_quadratic_is_exact(D, p):
    nodes = graded_nodes(D, 5)
    m = RefModel(D, None, p, 2 * p, nodes=nodes)
    nu = 0.7
    coef = [1.0, 2.0, -0.5][:D]
    u = lambda x: sum(c * x[:, d] ** 2 for d, c in enumerate(coef)) + 0.3 * x[:, 0] * x[:, 1]
    R = m.evaluate(OP_SIP, params_vector(1.3, nu), project(m, nodes, [u])).reshape(m.J, m.N)
    # R[j, n] = int_K nu lap(u) psi_n: only mode 0 is nonzero, nu lap(u) |K| scaling phi_0
    g = m.grid()
    expected0 = nu * 2.0 * sum(coef) * (2.0 ** D) * g["detJ"] * g["scaling"] * (2.0 ** (-D / 2.0))
    inner = interior(nodes)
    assert inner.sum() >= 4
    assert np.abs(R[inner, 0] - expected0[inner]).max() < 1e-10 * np.abs(expected0).max()
    assert np.abs(R[inner, 1:]).max() < 1e-10 * np.abs(expected0).max()


def test_graded_advection_diffusion_exact_on_quadratic():
    nodes = graded_nodes(3, 4, seed=7)
    m = RefModel(3, None, 2, 4, nodes=nodes)
    vel, nu = (1.0, 0.5, 0.25), -0.01
    u = lambda x: x[:, 0] ** 2 - 2 * x[:, 1] + 3 * x[:, 2] ** 2 + x[:, 0] * x[:, 2]
    grad = lambda x: np.stack([2 * x[:, 0] + x[:, 2], -2 * np.ones(len(x)), 6 * x[:, 2] + x[:, 0]], 1)
    R = m.evaluate(OP_ADVDIFF, params_vector(1.3, nu, vel), project(m, nodes, [u])).reshape(m.J, m.N)
    src = project(m, nodes, [lambda x: grad(x) @ np.array(vel) + nu * (2 + 6)]).reshape(m.J, m.N)
    inner = interior(nodes)
    assert np.abs(R[inner] - src[inner]).max() < 1e-11 * np.abs(src).max()


@pytest.mark.parametrize("kind", [OP_EULER_LF, OP_EULER_HLLC])
def test_graded_euler_constant_state_and_conservation(kind):
    D, p, gamma = 3, 2, 1.4
    nodes = graded_nodes(D, 3, seed=9)
    m = RefModel(D, None, p, 6, periodic=[1] * D, nodes=nodes)
    prm = params_vector(1.3, 1.0, (0, 0, 0), gamma)
    V = D + 2
    rho, vel, pr = 1.1, np.array([0.3, -0.2, 0.1]), 0.9
    state = [rho, *(rho * vel), pr / (gamma - 1) + 0.5 * rho * vel @ vel]
    U = project(m, nodes, [lambda x, c=c: c * np.ones(len(x)) for c in state])
    R = m.evaluate(kind, prm, U)
    assert np.abs(R).max() < 1e-12 * np.abs(U).max()
    # smooth non-constant state: sum over cells of the mean-mode residual / scaling vanishes
    U = project(m, nodes, [lambda x, c=c, k=k: c * (1 + 0.05 * np.sin(2 * np.pi * (x[:, 0] + k * x[:, 1] - x[:, 2])))
                           for k, c in enumerate(state)])
    R = m.evaluate(kind, prm, U).reshape(m.J, V, m.N)
    g = m.grid()
    tot = (R[:, :, 0] / g["scaling"][:, None]).sum(0)
    scale = (np.abs(R[:, :, 0]) / g["scaling"][:, None]).sum(0)
    assert np.abs(tot).max() < 1e-13 * scale.max()
    # the per-cell residual is far from zero, so the sum is a real cancellation
    assert scale.max() > 1e-3


@pytest.mark.parametrize("kind,D,p", [(OP_SIP, 2, 2), (OP_ADVDIFF, 3, 2), (OP_EULER_LF, 3, 2), (OP_EULER_HLLC, 3, 2),
                                      (OP_EULER_HLLC, 3, 3)])
def test_graded_jvp_vjp_against_fd_and_adjoint_identity(kind, D, p):
    """JVP against central differences of the residual (h = 1e-6); both VJP routines (the
    structured one and the per-cell dense one) against the JVP through <W, J dU> = <J^T W, dU>
    and against the transposed assembled Jacobian."""
    euler = kind in (OP_EULER_LF, OP_EULER_HLLC)
    nodes = graded_nodes(D, 3, seed=11)
    m = RefModel(D, None, p, 3 * p if euler else 2 * p, periodic=[int(kind != OP_SIP)] * D, nodes=nodes)
    prm = params_vector(1.3, -0.01 if kind == OP_ADVDIFF else 1.0, (1.0, 0.5, 0.25), 1.4)
    rng = np.random.default_rng(3)
    if euler:
        # piecewise-constant random state: the jumps across faces keep the two sides' wave
        # speeds apart (max/min of them are kinks central differences must not straddle), and
        # each velocity component has one sign, so u.n stays clear of the HLLC contact switch
        gamma, V, s0 = 1.4, D + 2, np.sqrt(m.grid()["detJ"]) * 2.0 ** (D / 2.0)
        rho, p_ = rng.uniform(0.8, 1.2, m.J), rng.uniform(0.9, 1.1, m.J)
        vel = rng.uniform(0.2, 0.4, (m.J, D)) * np.array([1.0, -1.0, 1.0])[:D]
        Uc = np.zeros((m.J, V, m.N))
        Uc[:, 0, 0] = rho * s0
        Uc[:, 1:1 + D, 0] = rho[:, None] * vel * s0[:, None]
        Uc[:, D + 1, 0] = (p_ / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)) * s0
        Uc[:, :, 1:] = 0.01 * rng.uniform(-1, 1, (m.J, V, m.N - 1)) * s0[:, None, None]
        U = Uc.reshape(-1)
    else:
        U = rng.uniform(-1, 1, m.J * m.N)
    dU = rng.uniform(-1, 1, U.shape) * np.abs(U).max() * 0.1
    W = rng.uniform(-1, 1, U.shape)
    Jd = m.jvp(kind, prm, U, dU)
    h = 1e-6
    fd = (m.evaluate(kind, prm, U + h * dU) - m.evaluate(kind, prm, U - h * dU)) / (2 * h)
    assert np.abs(Jd - fd).max() < 1e-7 * np.abs(fd).max()
    A = m.dense_matrix(kind, prm, U)
    assert np.abs(A @ dU - Jd).max() < 1e-12 * np.abs(Jd).max()
    for JtW in (m.vjp(kind, prm, U, W), m.vjp_dense(kind, prm, U, W)):
        lhs, rhs = W @ Jd, JtW @ dU
        assert abs(lhs - rhs) < 1e-12 * np.abs(W).sum() * np.abs(Jd).max()
        assert np.abs(A.T @ W - JtW).max() < 1e-12 * np.abs(JtW).max()
