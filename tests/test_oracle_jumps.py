"""Pins of the oracle's flux JUMP terms: the parts of the inner-edge forms that only act where
the two traces differ.  Continuous fields, constant states, conservation and derivative
identities cannot see them; piecewise-constant states can, and for those the mode-0 (and
x-linear mode) residuals have closed forms.

Set-up: a periodic row of three cells along one axis (one cell across the other axes, whose
self-periodic faces cancel).  Cells 0 and 1 hold the state UL, cell 2 holds UR.  The face
(0|1) carries F(UL).n by consistency, so for cell 1

    R[1, :, mode 0] = (Fhat(UL, UR, +axis) - F(UL).axis) * A / sqrt(|K_1|)

isolates ONE Riemann problem, which is compared with flux functions written here from their
textbook definitions, independently of the oracle's code:

  * HLLC: Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics" (3rd ed.), sections
    10.4-10.6: PVRS pressure estimate, the q-factor wave speeds S_L, S_R, the contact speed S*
    and the star states of eq. 10.73, fluxes F*_K = F_K + S_K (U*_K - U_K), the four-way choice
    of eq. 10.71.  The reference's form (HLLCFlux.cs:48-86, HLLC{Density,Momentum,Energy}Flux.cs,
    EulerFlux.cs:216-251) is non-dimensional; with Ma = 1/sqrt(gamma) (the reference tests'
    choice, ShockTubeTests.cs:333-334) gamma Ma^2 = 1 and it is the dimensional system.
  * Rusanov: 1/2 (F_L + F_R - S (U_R - U_L)), S = max(|u_L.n| + a_L, |u_R.n| + a_R).
  * upwind transport: F = (beta.n) u_upwind (TransportFlux.cs:157-164).
  * SIP interior penalty on graded cells: -mu nu [u][v] with
    mu = penalty * max((p+1)^2, (p+1)(p+D)/D) / min(LengthScale_A, LengthScale_B)
    (ipLaplace.cs:128-140, 194-216), LengthScale = 2 |K| / |dK| (GridData_CellData.cs:681-684),
    and the symmetry term 1/2 nu (grad v_A + grad v_B).n [u] on the x-linear test function.

Each planted mistake of the restated forms -- the HLLC energy star factor S* gamma Ma^2 doubled,
the downwind instead of the upwind value, the interior penalty doubled, the max of the two
cells' penalties replaced by the min -- fails at least one of these tests.
"""
import numpy as np
import pytest

from oracle.bosss_ref import RefModel, params_vector, OP_ADVDIFF, OP_EULER_HLLC, OP_EULER_LF, OP_SIP

GAMMA = 1.4


# ------------------------------------------------------------------ textbook fluxes
def prim_to_cons(rho, u, p, gamma=GAMMA):
    u = np.asarray(u, float)
    return np.concatenate([[rho], rho * u, [p / (gamma - 1) + 0.5 * rho * u @ u]])


def cons_to_prim(U, gamma=GAMMA):
    D = len(U) - 2
    rho, m, E = U[0], U[1:1 + D], U[1 + D]
    u = m / rho
    return rho, u, (gamma - 1) * (E - 0.5 * rho * u @ u)


def physical_flux(U, n, gamma=GAMMA):
    rho, u, p = cons_to_prim(U, gamma)
    un = u @ n
    return np.concatenate([[rho * un], rho * u * un + p * n, [un * (U[-1] + p)]])


def toro_hllc(UL, UR, n, gamma=GAMMA):
    """HLLC flux, Toro sections 10.4-10.6 (PVRS wave speeds, eq. 10.73 star states)."""
    rL, uL, pL = cons_to_prim(UL, gamma)
    rR, uR, pR = cons_to_prim(UR, gamma)
    aL, aR = np.sqrt(gamma * pL / rL), np.sqrt(gamma * pR / rR)
    unL, unR = uL @ n, uR @ n
    p_pvrs = 0.5 * (pL + pR) - 0.5 * (unR - unL) * 0.5 * (rL + rR) * 0.5 * (aL + aR)
    pstar = max(0.0, p_pvrs)

    def q(pK):
        return 1.0 if pstar <= pK else np.sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (pstar / pK - 1.0))

    SL, SR = unL - aL * q(pL), unR + aR * q(pR)
    Ss = (pR - pL + rL * unL * (SL - unL) - rR * unR * (SR - unR)) / (rL * (SL - unL) - rR * (SR - unR))

    def star(U, r, u, p, un, S):
        c = r * (S - un) / (S - Ss)
        return c * np.concatenate([[1.0], u + (Ss - un) * n, [U[-1] / r + (Ss - un) * (Ss + p / (r * (S - un)))]])

    FL, FR = physical_flux(UL, n, gamma), physical_flux(UR, n, gamma)
    if 0.0 <= SL:
        branch, F = "L", FL
    elif SL <= 0.0 <= Ss:
        branch, F = "*L", FL + SL * (star(UL, rL, uL, pL, unL, SL) - UL)
    elif Ss <= 0.0 <= SR:
        branch, F = "*R", FR + SR * (star(UR, rR, uR, pR, unR, SR) - UR)
    else:
        branch, F = "R", FR
    return F, dict(branch=branch, SL=SL, SR=SR, Sstar=Ss, pstar=pstar, shockL=pstar > pL, shockR=pstar > pR)


def rusanov(UL, UR, n, gamma=GAMMA):
    rL, uL, pL = cons_to_prim(UL, gamma)
    rR, uR, pR = cons_to_prim(UR, gamma)
    S = max(abs(uL @ n) + np.sqrt(gamma * pL / rL), abs(uR @ n) + np.sqrt(gamma * pR / rR))
    return 0.5 * (physical_flux(UL, n, gamma) + physical_flux(UR, n, gamma) - S * (UR - UL))


# ------------------------------------------------------------------ three-cell row
def row_model(D, axis, p, order, nodes_along=None):
    """Periodic grid of 3 cells along `axis` and 1 across the others, on the unit box (or on
    the given nodes along `axis`)."""
    nodes = [np.array([0.0, 1.0]) for _ in range(D)]
    nodes[axis] = np.array([0.0, 1 / 3, 2 / 3, 1.0]) if nodes_along is None else np.asarray(nodes_along, float)
    m = RefModel(D, None, p, order, periodic=[1] * D, nodes=nodes)
    widths = np.diff(nodes[axis])
    return m, widths


def oracle_face_flux(kind, UL, UR, axis, D=3, p=1):
    """Fhat(UL, UR, +axis) as the oracle applies it, recovered from cell 1's mode-0 residual."""
    m, w = row_model(D, axis, p, 2 * p + 2)
    V = D + 2
    vol = w[1]                                   # |K| of cell 1 (unit extent across the axis)
    sqrtK = np.sqrt(vol)
    U = np.zeros((3, V, m.N))
    for j, s in enumerate((UL, UL, UR)):
        U[j, :, 0] = s * np.sqrt(w[j])           # constant state: coefficient = value * sqrt(|K|)
    R = m.evaluate(kind, params_vector(gamma=GAMMA), U.reshape(-1)).reshape(3, V, m.N)
    n = np.eye(D)[axis]
    return R[1, :, 0] * sqrtK / 1.0 + physical_flux(UL, n)   # face area A = 1


# states: (rho, velocity, p) for L and R, and the HLLC branch the textbook flux takes
HLLC_CASES = {
    "subsonic_contact_right": ((1.0, (0.30, 0.10, -0.05), 1.0), (0.8, (0.10, -0.20, 0.15), 0.9), "*L"),
    "subsonic_contact_left": ((1.0, (-0.10, 0.10, 0.02), 1.0), (0.9, (-0.30, 0.05, -0.10), 1.1), "*R"),
    "two_shocks": ((1.1, (0.60, -0.10, 0.20), 1.0), (0.9, (-0.50, 0.30, 0.00), 0.95), None),
    "two_rarefactions_vacuum_pvrs": ((1.0, (-2.5, 0.20, 0.10), 0.4), (1.0, (2.5, -0.10, 0.30), 0.4), None),
    "supersonic_right": ((1.0, (2.0, 0.10, 0.0), 1.0), (0.9, (2.2, -0.1, 0.05), 0.95), "L"),
    "supersonic_left": ((1.0, (-2.2, 0.10, 0.0), 1.0), (0.9, (-2.0, -0.1, 0.05), 0.95), "R"),
}


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("case", sorted(HLLC_CASES))
def test_hllc_jump_matches_toro(case, axis):
    (rl, ul, pl), (rr, ur, pr), branch = HLLC_CASES[case]
    # rotate the velocities so that the normal component lies along `axis`
    perm = [(axis + k) % 3 for k in range(3)]
    vl, vr = np.zeros(3), np.zeros(3)
    vl[perm], vr[perm] = ul, ur
    UL, UR = prim_to_cons(rl, vl, pl), prim_to_cons(rr, vr, pr)
    n = np.eye(3)[axis]
    F, info = toro_hllc(UL, UR, n)
    if branch is not None:
        assert info["branch"] == branch
    if case == "two_shocks":
        assert info["shockL"] and info["shockR"] and info["branch"] in ("*L", "*R")
    if case == "two_rarefactions_vacuum_pvrs":
        assert info["pstar"] == 0.0 and info["branch"] in ("*L", "*R")
    got = oracle_face_flux(OP_EULER_HLLC, UL, UR, axis)
    assert np.abs(got - F).max() < 1e-12 * np.abs(F).max(), (case, got, F, info)


@pytest.mark.parametrize("axis", [0, 2])
@pytest.mark.parametrize("case", ["subsonic_contact_right", "subsonic_contact_left", "two_shocks"])
def test_rusanov_jump_matches_closed_form(case, axis):
    (rl, ul, pl), (rr, ur, pr), _ = HLLC_CASES[case]
    perm = [(axis + k) % 3 for k in range(3)]
    vl, vr = np.zeros(3), np.zeros(3)
    vl[perm], vr[perm] = ul, ur
    UL, UR = prim_to_cons(rl, vl, pl), prim_to_cons(rr, vr, pr)
    F = rusanov(UL, UR, np.eye(3)[axis])
    got = oracle_face_flux(OP_EULER_LF, UL, UR, axis)
    assert np.abs(got - F).max() < 1e-12 * np.abs(F).max()


def test_hllc_differs_from_rusanov_on_these_states():
    # guards the test above against states on which the two fluxes coincide
    (rl, ul, pl), (rr, ur, pr), _ = HLLC_CASES["subsonic_contact_right"]
    UL, UR = prim_to_cons(rl, ul, pl), prim_to_cons(rr, ur, pr)
    n = np.eye(3)[0]
    assert np.abs(toro_hllc(UL, UR, n)[0] - rusanov(UL, UR, n)).max() > 1e-3


# ------------------------------------------------------------------ upwind transport
@pytest.mark.parametrize("vx", [1.0, -1.0])
def test_upwind_takes_the_upwind_value(vx):
    # pure transport (nu = 0): cells 0, 1 hold uL, cell 2 holds uR; velocity (vx, 0.5, 0.25)
    D, p = 3, 1
    m, w = row_model(D, 0, p, 2 * p)
    uL, uR = 0.7, -0.4
    vals = (uL, uL, uR)
    U = np.zeros((3, m.N))
    for j in range(3):
        U[j, 0] = vals[j] * np.sqrt(w[j])
    R = m.evaluate(OP_ADVDIFF, params_vector(1.3, 0.0, (vx, 0.5, 0.25)), U.reshape(-1)).reshape(3, m.N)
    # face k+1/2 between cells k, k+1 (periodic): flux (beta.x) * (value of the upwind cell)
    face = [vx * (vals[k] if vx > 0 else vals[(k + 1) % 3]) for k in range(3)]
    for j in range(3):
        expect = (face[j] - face[j - 1]) / np.sqrt(w[j])        # out through +x, in through -x
        assert abs(R[j, 0] - expect) < 1e-13, (j, R[j, 0], expect)
    assert abs(R[1, 0]) < 1e-13 if vx > 0 else abs(R[1, 0]) > 0.1


# ------------------------------------------------------------------ SIP interior penalty
def _x_mode(m):
    """Index and scale c of the mode phi_n = c * xi_x (reference coordinates)."""
    t = m.tables()
    xi = t["vnodes"][:, 0]
    for n in range(m.N):
        c = t["phi"][0, n] / xi[0] if xi[0] != 0 else None
        if c is not None and np.allclose(t["phi"][:, n], c * xi, atol=1e-14):
            return n, c
    raise AssertionError("no x-linear mode")


@pytest.mark.parametrize("nu", [1.0, -0.01])
@pytest.mark.parametrize("p", [1, 2])
def test_sip_interior_penalty_on_graded_cells(p, nu):
    D, pen = 3, 1.3
    xs = [0.0, 0.2, 0.5, 1.0]                                     # widths 0.2, 0.3, 0.5
    m, h = row_model(D, 0, p, 2 * p, nodes_along=xs)
    vals = np.array([0.9, -0.3, 0.5])                             # a different value in every cell
    U = np.zeros((3, m.N))
    for j in range(3):
        U[j, 0] = vals[j] * np.sqrt(h[j])
    R = m.evaluate(OP_SIP, params_vector(pen, nu), U.reshape(-1)).reshape(3, m.N)
    vol, surf = h * 1.0, 2.0 * (1.0 + 2.0 * h)                    # cells h x 1 x 1, all faces inner
    ls = 2.0 * vol / surf                                         # CellLengthScale
    pdeg = max((p + 1.0) ** 2, (p + 1.0) * (p + D) / D)

    def mu(a, b):                                                 # penalty of the face between a, b
        return pen * pdeg / min(ls[a], ls[b])

    nx, c = _x_mode(m)
    for j in range(3):
        l, r = (j - 1) % 3, (j + 1) % 3
        # mode 0: only the penalty sees a jump of constants (grad u = 0, grad v = 0)
        exp0 = -nu / np.sqrt(h[j]) * (mu(j, r) * (vals[j] - vals[r]) + mu(l, j) * (vals[j] - vals[l]))
        assert abs(R[j, 0] - exp0) < 1e-12 * max(1.0, abs(exp0)), (j, R[j, 0], exp0)
        # x-linear mode v = c sc xi_x: trace +-c sc at the +x / -x face, d v/dx = 2 c sc / h
        sc = 1.0 / np.sqrt(h[j] / 8.0)
        cv, g = c * sc, 2.0 * c * sc / h[j]
        expx = nu * (0.5 * g * (vals[j] - vals[r]) - mu(j, r) * (vals[j] - vals[r]) * cv
                     + 0.5 * g * (vals[l] - vals[j]) - mu(l, j) * (vals[l] - vals[j]) * cv)
        assert abs(R[j, nx] - expx) < 1e-12 * max(1.0, abs(expx)), (j, R[j, nx], expx)
    assert len(set(np.round(ls, 12))) == 3                       # the penalty rule sees unequal cells
