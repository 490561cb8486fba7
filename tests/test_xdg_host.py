"""XDG Poisson: cut-cell quadrature (host) and the oracle's XDG operator, pinned without a GPU.

Pins:
  * cut-cell rules: sphere volume (4/3) pi r^3 and area 4 pi r^2, circle area / perimeter
    (closed forms, cf. CutCellQuadrature/TestCases/TestCases3D.cs sphere cases); species
    volumes of every cell and species areas of every face add up to the full measures;
    a planar interface gives the face area exactly;
  * species length scales of uncut cells: |K| / |dK| (CutCellMetrics + LengthScaleAgg);
  * the oracle's XDG operator: the reference's own known-answer test, XdgPoisson3
    ParabolaTest (Tests.cs:80-104, HardCodedControl.PiecewiseParabola: 2D, planar interface,
    mu_A = -20, mu_B = -1, agglomeration threshold 0, p = 2): the piecewise quadratic exact
    solution is reproduced to L2 error < 1e-10; for mu_A = mu_B and a continuous quadratic
    the species-summed XDG residual equals the standard SIP residual of the restated
    single-phase evaluator on every cell away from the boundary (interface terms cancel,
    cut-cell rules integrate the split polynomials exactly); the matrix is symmetric.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from bosss_amd.grid import CartesianGrid
from bosss_amd.refelement import basis, num_modes
from bosss_amd.xdg import QuadLevelSet, XdgRules, tensor_nodes
from oracle.bosss_ref import RefModel, OP_SIP, params_vector, xdg_assemble, xdg_to_csr


@pytest.mark.parametrize("n1,n,tol", [(8, 8, 1e-8), (16, 8, 1e-13)])
def test_sphere_volume_and_area(n1, n, tol):
    r = 0.3
    R = XdgRules(CartesianGrid((n1,) * 3), QuadLevelSet.sphere([0.5] * 3, r), n=n)
    assert abs(R.volumes[:, 0].sum() / (4 / 3 * np.pi * r ** 3) - 1) < tol
    assert abs(R.interface_area.sum() / (4 * np.pi * r ** 2) - 1) < tol
    assert np.abs(R.volumes.sum(1) - R.grid.volume).max() < 1e-15
    # species parts of every face add up to the face
    g = R.grid
    farea = g.sqrtgram * 2 ** (g.D - 1)
    assert np.abs(R.edge_area.sum(1) - farea).max() < 1e-15


def test_circle_off_grid_center():
    # centre not on a grid line: exercises the split at the vertex of the quadratic
    r = 0.27
    R = XdgRules(CartesianGrid((9, 9)), QuadLevelSet.sphere([0.47, 0.52], r), n=10)
    assert abs(R.volumes[:, 0].sum() / (np.pi * r * r) - 1) < 1e-12
    assert abs(R.interface_area.sum() / (2 * np.pi * r) - 1) < 1e-12
    assert np.abs(R.volumes.sum(1) - R.grid.volume).max() < 1e-15


def test_plane_exact_and_lengthscales():
    g = CartesianGrid((4, 3, 3))
    R = XdgRules(g, QuadLevelSet.plane(3, 0, 0.4), n=3)
    assert abs(R.interface_area.sum() - 1.0) < 1e-14
    assert abs(R.volumes[:, 0].sum() - 0.4) < 1e-14
    unc = ~R.cut
    vol = g.volume[unc]
    surf = 2 * (g.hc[unc, 0] * g.hc[unc, 1] + g.hc[unc, 1] * g.hc[unc, 2] + g.hc[unc, 0] * g.hc[unc, 2])
    sp = np.where(R.has[unc, 0], 0, 1)
    assert np.abs(R.lenscale[unc][np.arange(unc.sum()), sp] - vol / surf).max() < 1e-15
    # cut cells: LS_s = Vol_s / (interface + species face areas)
    j = np.where(R.cut)[0][0]
    fa = R.edge_area[(g.cells[:, 0] == j) | (g.cells[:, 1] == j)].sum(0)
    assert np.allclose(R.lenscale[j], R.volumes[j] / (R.interface_area[j] + fa), rtol=1e-14)


def parabola_case(delta, muA=-20.0, muB=-1.0, p=2):
    """HardCodedControl.PiecewiseParabola (HardCodedControl.cs:314-375)."""
    xn = np.array([-6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6.0])
    g = CartesianGrid(nodes=[xn, np.linspace(-3, 3, 7)], periodic=[False, True])
    R = XdgRules(g, QuadLevelSet.plane(2, 0, delta), n=3)
    A0 = 36 / (muB + muA); A1 = 3 * (-muB + muA) / (muA * (muB + muA)); A2 = -1 / (2 * muA)
    B0 = 36 / (muB + muA); B1 = 3 * (-muB + muA) / ((muB + muA) * muB); B2 = -1 / (2 * muB)
    uA = lambda X: -(A2 * (X[:, 0] - delta) ** 2 + A1 * (X[:, 0] - delta) + A0)
    uB = lambda X: -(B2 * (X[:, 0] - delta) ** 2 + B1 * (X[:, 0] - delta) + B0)
    dl, dr = uA(np.array([[-6.0, 0]]))[0], uB(np.array([[6.0, 0]]))[0]
    g_D = lambda X: np.where(X[:, 0] < 0, dl, dr)
    return R, (muA, muB), (uA, uB), g_D


def xdg_l2_error(R, p, u, exact):
    """L2 error over both species domains with the XDG volume rules."""
    N = num_modes(R.D, p)
    lay = R.block_layout(N)
    g = R.grid
    err = 0.0
    for it in range(len(R.vol_cell)):
        j, sp = R.vol_cell[it], R.vol_spec[it]
        a, b = R.vol_off[it], R.vol_off[it + 1]
        xi = R.vol_nodes[a:b]
        phi = basis(R.D, p, xi)[0] * g.scaling[j]
        o = lay["dof"][j] + R.species_offset(j, sp, N)
        uh = phi @ u[o:o + N]
        X = g.centers[j] + 0.5 * g.hc[j] * xi
        err += (R.vol_w[a:b] * (uh - exact[sp](X)) ** 2).sum()
    return np.sqrt(err)


@pytest.mark.parametrize("delta", [0.0, 0.3])
def test_oracle_parabola_known_answer(delta):
    R, mu, ex, g_D = parabola_case(delta)
    blocks, aff, lay = xdg_assemble(R, 2, mu, [1.0, 1.0], penalty=2.0, gD=R.boundary_values(g_D))
    M = xdg_to_csr(R, lay, blocks)
    u = spla.spsolve(M.tocsc(), -aff)
    assert xdg_l2_error(R, 2, u, ex) < 1e-10
    assert abs(M - M.T).max() <= 1e-12 * abs(M).max()


def test_oracle_species_sum_equals_single_phase_sip():
    n, p, mu = 6, 2, -1.3
    g = CartesianGrid((n,) * 3)
    R = XdgRules(g, QuadLevelSet.sphere([0.5] * 3, 0.3), n=3)
    blocks, aff, lay = xdg_assemble(R, p, [mu, mu], [0.0, 0.0], penalty=2.0)
    M = xdg_to_csr(R, lay, blocks)
    N = num_modes(3, p)
    # continuous quadratic, projected exactly on every cell
    q = lambda X: 1.0 + X[:, 0] - 2 * X[:, 1] * X[:, 2] + 0.5 * X[:, 2] ** 2 + X[:, 0] * X[:, 1]
    xr, wr = tensor_nodes(-np.ones(3), np.ones(3), 4)
    phi = basis(3, p, xr)[0]
    Ustd = np.zeros(g.J * N)
    for j in range(g.J):
        Ustd[j * N:(j + 1) * N] = (phi * (wr * q(g.centers[j] + 0.5 * g.hc[j] * xr))[:, None]).sum(0) / g.scaling[j]
    Ux = np.zeros(lay["ndof"])
    for j in range(g.J):
        for sp in (0, 1):
            o = R.species_offset(j, sp, N)
            if o >= 0:
                Ux[lay["dof"][j] + o: lay["dof"][j] + o + N] = Ustd[j * N:(j + 1) * N]
    Rx = M @ Ux + aff
    ref = RefModel(3, (n,) * 3, p, 2 * p)
    Rs = ref.evaluate(OP_SIP, params_vector(2.0, mu), Ustd)
    idx = np.indices((n,) * 3).reshape(3, -1).T
    interior = np.all((idx > 0) & (idx < n - 1), axis=1)
    worst = 0.0
    for j in np.where(interior)[0]:
        s = np.zeros(N)
        for sp in (0, 1):
            o = R.species_offset(j, sp, N)
            if o >= 0:
                s += Rx[lay["dof"][j] + o: lay["dof"][j] + o + N]
        worst = max(worst, np.abs(s - Rs[j * N:(j + 1) * N]).max())
    assert R.cut[interior].sum() > 20
    assert worst < 1e-11 * np.abs(Rs).max()

