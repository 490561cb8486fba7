"""Pins of the oracle's DG operator against closed forms and invariants.

Restates the reference's own operator checks: DerivativeTestMain.cs:694-728,
965-1176 (exact Laplacian of polynomials, Evaluate == M*U + b, FD Jacobian vs
operator matrix), plus consistency/conservation invariants of the Euler fluxes
(EulerFlux.cs, HLLCFlux.cs:48-86, RusanovFlux.cs:55-65).
"""
import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_SIP, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC


def project(m, funcs):
    """L2-projection of point functions onto the DG basis (orthonormal => c = int f psi)."""
    g, t = m.grid(), m.tables()
    D = m.D
    h = np.array([1.0 / n for n in m.ncells])
    centers = _centers(m)
    V = len(funcs)
    U = np.zeros((m.J, V, m.N))
    for j in range(m.J):
        x = centers[j] + 0.5 * h * t["vnodes"]
        for f, fn in enumerate(funcs):
            U[j, f] = (t["vw"] * fn(x) * g["detJ"][j] * g["scaling"][j]) @ t["phi"]
    return U.reshape(-1)


def _centers(m):
    D = m.D
    idx = np.indices(m.ncells).reshape(D, -1).T  # i slowest ... last index fastest
    h = np.array([1.0 / n for n in m.ncells])
    return (idx + 0.5) * h


def make(D, n, p, order, periodic=False):
    m = RefModel(D, (n,) * D, p, order, periodic=[int(periodic)] * D)
    m.ncells = (n,) * D
    return m


def interior_cells(m, layers=1):
    idx = np.indices(m.ncells).reshape(m.D, -1).T
    n = np.array(m.ncells)
    return np.all((idx >= layers) & (idx < n - layers), axis=1)


@pytest.mark.parametrize("D,p", [(2, 2), (3, 2), (3, 3)])
def test_sip_laplacian_of_quadratic_is_exact(D, p):
    m = make(D, 5, p, 2 * p)
    nu = 0.7
    prm = params_vector(1.3, nu)
    coef = [1.0, 2.0, -0.5][:D]
    u = lambda x: sum(c * x[:, d] ** 2 for d, c in enumerate(coef)) + 0.3 * x[:, 0] * x[:, 1]
    U = project(m, [u])
    R = m.evaluate(OP_SIP, prm, U).reshape(m.J, m.N)
    lap = 2.0 * sum(coef)
    g = m.grid()
    # R[j,n] = int nu*lap*psi_n ; int psi_0 = |K| * Scaling * phi_0
    expected0 = nu * lap * (2.0 ** D) * g["detJ"] * g["scaling"] * (2.0 ** (-D / 2.0))
    inner = interior_cells(m)
    assert np.abs(R[inner, 0] - expected0[inner]).max() < 1e-10 * np.abs(expected0).max()
    assert np.abs(R[inner, 1:]).max() < 1e-10 * np.abs(expected0).max()


def test_sip_single_cell_dirichlet_closed_form():
    # one cell [0,1]^2, four Dirichlet edges, u = 1:
    # R_0 = sum_edges int (-2 mu psi_0) ds, psi_0 = 1, mu = 1.3 * 9 / lenscale, lenscale = 2|K|/(4+4)
    m = make(2, 1, 2, 4)
    R = m.evaluate(OP_SIP, params_vector(1.3, 1.0), project(m, [lambda x: np.ones(len(x))]))
    assert abs(m.grid()["lenscale"][0] - 0.25) < 1e-15
    mu = 1.3 * 9 / 0.25
    assert abs(R[0] - (-2 * mu * 4.0)) < 1e-11
    assert np.abs(R[1:3]).max() < 1e-11  # odd modes cancel between opposite edges


@pytest.mark.parametrize("D,p", [(2, 2), (3, 1), (3, 2)])
def test_sip_matrix_symmetric_negative_definite_and_linear(D, p):
    m = make(D, 3, p, 2 * p)
    prm = params_vector(1.3, 1.0)
    A = m.dense_matrix(OP_SIP, prm)
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 0
    U = np.random.default_rng(1).uniform(-1, 1, m.J * m.N)
    R = m.evaluate(OP_SIP, prm, U)
    assert np.abs(R - A @ U).max() < 1e-12 * np.abs(R).max()
    # alpha/beta semantics: R <- beta R + alpha op(U)
    R0 = np.random.default_rng(2).uniform(-1, 1, U.shape)
    R2 = m.evaluate(OP_SIP, prm, U, alpha=0.5, beta=2.0, R=R0.copy())
    assert np.abs(R2 - (2.0 * R0 + 0.5 * A @ U)).max() < 1e-12 * np.abs(R2).max()


def test_advection_diffusion_exact_on_quadratic():
    m = make(3, 4, 2, 4)
    vel = (1.0, 0.5, 0.25)
    nu = -0.01
    prm = params_vector(1.3, nu, vel)
    u = lambda x: x[:, 0] ** 2 - 2 * x[:, 1] + 3 * x[:, 2] ** 2 + x[:, 0] * x[:, 2]
    # div(beta u) + nu lap u
    gradlin = lambda x: np.stack([2 * x[:, 0] + x[:, 2], -2 * np.ones(len(x)), 6 * x[:, 2] + x[:, 0]], 1)
    U = project(m, [u])
    R = m.evaluate(OP_ADVDIFF, prm, U).reshape(m.J, m.N)
    src = project(m, [lambda x: gradlin(x) @ np.array(vel) + nu * (2 + 6)]).reshape(m.J, m.N)
    inner = interior_cells(m)
    assert np.abs(R[inner] - src[inner]).max() < 1e-11 * np.abs(src).max()


def test_advection_constant_periodic_is_zero():
    m = make(3, 3, 3, 6, periodic=True)
    U = project(m, [lambda x: 2.5 * np.ones(len(x))])
    R = m.evaluate(OP_ADVDIFF, params_vector(1.3, -0.01, (1.0, 0.5, 0.25)), U)
    assert np.abs(R).max() < 1e-12


def _euler_state(m, seed=3, smooth=True):
    rng = np.random.default_rng(seed)
    D = m.D
    V = D + 2
    gamma = 1.4
    U = np.zeros((m.J, V, m.N))
    rho = rng.uniform(0.8, 1.2, m.J)
    vel = rng.uniform(-0.1, 0.1, (m.J, D))
    p = rng.uniform(0.9, 1.1, m.J)
    E = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)
    s0 = m.grid()["detJ"] ** 0.5 * 2.0 ** (D / 2.0)  # mode-0 coefficient of a constant 1
    U[:, 0, 0] = rho * s0
    U[:, 1:1 + D, 0] = (rho[:, None] * vel) * s0[:, None]
    U[:, D + 1, 0] = E * s0
    if smooth:
        U[:, :, 1:] = 0.01 * rng.uniform(-1, 1, (m.J, V, m.N - 1)) * s0[:, None, None]
    return U.reshape(-1)


@pytest.mark.parametrize("kind", [OP_EULER_LF, OP_EULER_HLLC])
@pytest.mark.parametrize("D", [2, 3])
def test_euler_constant_state_periodic_is_zero(kind, D):
    m = make(D, 3, 2, 6, periodic=True)
    U = _euler_state(m, smooth=False).reshape(m.J, D + 2, m.N)
    U[:, :, 0] = U[0, :, 0]
    R = m.evaluate(kind, params_vector(gamma=1.4), U.reshape(-1))
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("kind", [OP_EULER_LF, OP_EULER_HLLC])
def test_euler_conservation_periodic(kind):
    m = make(3, 3, 2, 6, periodic=True)
    U = _euler_state(m)
    R = m.evaluate(kind, params_vector(gamma=1.4), U).reshape(m.J, 5, m.N)
    # sum_j int R psi_0 = 0  (OUT flux = -IN flux, NECQuadratureEdge.cs:1000)
    assert np.abs(R[:, :, 0].sum(0)).max() < 1e-12 * np.abs(R).max()


def _primitive_flux(Uc, n, gamma, D):
    rho, m, E = Uc[0], Uc[1:1 + D], Uc[1 + D]
    u = m / rho
    p = (gamma - 1) * (E - 0.5 * rho * u @ u)
    un = u @ n
    return np.concatenate([[rho * un], m * un + p * n, [un * (E + p)]]), un, np.sqrt(gamma * p / rho)


@pytest.mark.parametrize("kind", [OP_EULER_LF, OP_EULER_HLLC])
def test_euler_piecewise_constant_face_fluxes(kind):
    # two cells in x (periodic), piecewise-constant states; mode-0 residual is the face-flux balance
    m = RefModel(2, (2, 1), 2, 6, periodic=[1, 1])
    gamma = 1.4
    D = 2
    s0 = m.grid()["detJ"][0] ** 0.5 * 2.0
    UL = np.array([1.0, 2.5, 0.1, 0.0]); UL[3] = 1.0 / (gamma - 1) + 0.5 * (2.5 ** 2 + 0.1 ** 2)  # supersonic in +x
    UR = np.array([0.9, 2.2, -0.1, 0.0]); UR[3] = 0.95 / (gamma - 1) + 0.5 * (2.2 ** 2 + 0.1 ** 2) / 0.9
    U = np.zeros((2, 4, m.N)); U[0, :, 0] = UL * s0; U[1, :, 0] = UR * s0
    R = m.evaluate(kind, params_vector(gamma=gamma), U.reshape(-1)).reshape(2, 4, m.N)
    n = np.array([1.0, 0.0])
    fL, unL, aL = _primitive_flux(UL, n, gamma, D)
    fR, unR, aR = _primitive_flux(UR, n, gamma, D)
    assert unL - aL > 0 and unR - aR > 0  # both supersonic to the right
    if kind == OP_EULER_HLLC:
        # supersonic: HLLC is the upwind flux F(U_in).n on the face L->R, F(U_R).n on the periodic face R->L
        f_LR, f_RL = fL, fR
    else:
        S = max(abs(unL) + aL, abs(unR) + aR)
        f_LR = 0.5 * (fL + fR - S * (UR - UL))
        f_RL = 0.5 * (fR + fL - S * (UL - UR))
    face = 1.0 * 0.5  # face length 1 (y-extent), psi_0 = 1/sqrt(|K|)
    psi0 = 1.0 / np.sqrt(0.5)
    # cell 0: +x face to cell 1 (flux f_LR out), -x face periodic from cell 1 (flux f_RL in)
    expect0 = psi0 * 1.0 * (f_LR - f_RL)
    assert np.abs(R[0, :, 0] - expect0).max() < 1e-12 * np.abs(expect0).max()
    assert np.abs(R[1, :, 0] + expect0).max() < 1e-12 * np.abs(expect0).max()
    del face


@pytest.mark.parametrize("kind,D,p", [(OP_SIP, 2, 2), (OP_ADVDIFF, 3, 2), (OP_EULER_LF, 3, 2), (OP_EULER_HLLC, 3, 2), (OP_EULER_HLLC, 2, 3)])
def test_jvp_vjp_against_fd_and_adjoint_identity(kind, D, p):
    periodic = kind in (OP_EULER_LF, OP_EULER_HLLC, OP_ADVDIFF)
    m = make(D, 3, p, 3 * p if kind >= OP_EULER_LF else 2 * p, periodic=periodic)
    prm = params_vector(1.3, -0.01 if kind == OP_ADVDIFF else 1.0, (1.0, 0.5, 0.25))
    rng = np.random.default_rng(7)
    V = m.nvars(kind)
    U = _euler_state(m) if kind >= OP_EULER_LF else rng.uniform(-1, 1, m.J * m.N)
    scale = np.abs(U).max()
    dU = rng.uniform(-1, 1, U.shape) * scale * 0.1
    W = rng.uniform(-1, 1, U.shape)
    Jd = m.jvp(kind, prm, U, dU)
    h = 1e-6
    fd = (m.evaluate(kind, prm, U + h * dU) - m.evaluate(kind, prm, U - h * dU)) / (2 * h)
    assert np.abs(Jd - fd).max() < 1e-7 * np.abs(fd).max()
    JtW = m.vjp(kind, prm, U, W)
    lhs, rhs = W @ Jd, JtW @ dU
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), np.abs(W).sum() * np.abs(Jd).max() * 1e-3)
    # assembled Jacobian (transpose) vs VJP
    A = m.dense_matrix(kind, prm, U)
    assert np.abs(A.T @ W - JtW).max() < 1e-12 * np.abs(JtW).max()
    assert np.abs(A @ dU - Jd).max() < 1e-12 * np.abs(Jd).max()
    assert V * m.N * m.J == U.size


def test_fd_jacobian_like_reference_matches_assembled():
    # FDJacobianBuilder: eps = sqrt(machine eps), relative with floor 1 (DifferentialOperator.cs:2621-2629)
    m = make(2, 2, 1, 2, periodic=True)
    kind = OP_EULER_HLLC
    prm = params_vector(gamma=1.4)
    U = _euler_state(m)
    R0 = m.evaluate(kind, prm, U)
    eps = np.sqrt(np.finfo(float).eps)
    A_fd = np.zeros((U.size, U.size))
    for i in range(U.size):
        d = eps * max(abs(U[i]), 1.0)
        Up = U.copy(); Up[i] += d
        A_fd[:, i] = (m.evaluate(kind, prm, Up) - R0) / d
    A = m.dense_matrix(kind, prm, U)
    assert np.abs(A - A_fd).max() < np.finfo(float).eps ** 0.25 * max(1, np.abs(A).max()) * 1e-2
