"""General polynomial level sets and Saye cut-cell rules for them (bosss_amd/levelset.py), pinned
by closed forms on the CPU:

  * the polynomial algebra: evaluation, gradient, restriction and line polynomials against direct
    evaluation; Bernstein bounds enclose sampled values on random boxes and are exact at corners;
  * the separable sphere as a tensor polynomial gives the same cells, species volumes, face areas
    and interface areas as the closed-form quadratic path;
  * a rotated ellipsoid (cross terms): volume (4/3) pi a b c, and the divergence identity
    int_Gamma (x - c).n dS = 3 |Omega| that pins the interface weights and normals together;
  * a torus (degree 4 in each variable): volume 2 pi^2 R r^2 and area 4 pi^2 R r;
  * species volumes of every cell and species areas of every face add up to the full measures.
"""
import numpy as np
import pytest

from bosss_amd.grid import CartesianGrid
from bosss_amd.levelset import PolyLevelSet
from bosss_amd.xdg import QuadLevelSet, XdgRules


def rotation(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def test_polynomial_algebra_and_bernstein_bounds():
    rng = np.random.default_rng(0)
    P = PolyLevelSet(rng.uniform(-1, 1, (4, 3, 5)))
    X = rng.uniform(-1, 2, (40, 3))
    direct = np.array([sum(P.C[i, j, k] * x[0] ** i * x[1] ** j * x[2] ** k for i in range(4) for j in range(3) for k in range(5))
                       for x in X])
    assert np.abs(P(X) - direct).max() < 1e-12 * np.abs(direct).max()
    h = 1e-6
    for d in range(3):
        e = np.zeros(3); e[d] = h
        fd = (P(X + e) - P(X - e)) / (2 * h)
        assert np.abs(P.grad(X)[:, d] - fd).max() < 1e-6 * np.abs(fd).max()
        Xr = X.copy(); Xr[:, d] = 0.7
        assert np.abs(P.restrict(d, 0.7)(np.delete(X, d, 1)) - P(Xr)).max() < 1e-12 * np.abs(P(Xr)).max()
        C = P.line_coeffs(d, np.delete(X, d, 1))
        assert np.abs(np.polynomial.polynomial.polyval(X[:, d], C.T, tensor=False) - direct).max() < 1e-12 * np.abs(direct).max()
    lo = rng.uniform(-1, 1, (20, 3))
    hi = lo + rng.uniform(0.05, 0.8, (20, 3))
    mn, mx = P.bounds(lo, hi)
    for b in range(20):
        S = lo[b] + rng.uniform(size=(500, 3)) * (hi[b] - lo[b])
        v = P(S)
        assert mn[b] <= v.min() + 1e-13 and v.max() <= mx[b] + 1e-13
        corners = np.array([[lo[b][d] if (c >> d) & 1 == 0 else hi[b][d] for d in range(3)] for c in range(8)])
        Bc = P.bernstein(lo[b], hi[b])[0]
        idx = [tuple(0 if (c >> d) & 1 == 0 else Bc.shape[d] - 1 for d in range(3)) for c in range(8)]
        assert np.abs(np.array([Bc[i] for i in idx]) - P(corners)).max() < 1e-12 * np.abs(Bc).max()


def test_polynomial_sphere_equals_quadratic_path():
    q = QuadLevelSet.sphere([0.47, 0.52, 0.5], 0.3)
    g = CartesianGrid((8, 8, 8))
    Rq = XdgRules(g, q, n=8)                          # 8 points per piece: both paths at rounding level
    Rp = XdgRules(g, PolyLevelSet.from_quadratic(q), n=8)
    assert (Rq.cut == Rp.cut).all() and (Rq.nspec == Rp.nspec).all() and (Rq.spec0 == Rp.spec0).all()
    # the paths split non-monotone boxes differently (at the vertex vs in halves): they agree to the
    # quadrature accuracy of 8 points per piece on 8^3 cells (faces: circle arcs near tangency)
    assert np.abs(Rq.volumes - Rp.volumes).max() < 1e-9
    assert np.abs(Rq.edge_area - Rp.edge_area).max() < 1e-6
    assert np.abs(Rq.interface_area - Rp.interface_area).max() < 1e-6


def test_rotated_ellipsoid_volume_and_divergence_identity():
    a = np.array([0.3, 0.2, 0.15])
    R = rotation(0.3, 0.5, 0.7)
    c = np.array([0.5, 0.47, 0.52])
    psi = PolyLevelSet.quadratic_form(R @ np.diag(1 / a ** 2) @ R.T, c)
    r = XdgRules(CartesianGrid((12, 12, 12)), psi, n=8)
    V = r.volumes[:, 0].sum()
    assert abs(V / (4 / 3 * np.pi * a.prod()) - 1) < 1e-7          # quadrature error at 12^3, 8 points per piece
    # interface nodes in reference coordinates -> physical: x = ctr + hc/2 xi
    g = r.grid
    cells = np.repeat(r.ifc_cell, np.diff(r.ifc_off))
    X = g.centers[cells] + 0.5 * g.hc[cells] * r.ifc_nodes
    flux = (r.ifc_w * ((X - c) * r.ifc_normal).sum(1)).sum()
    assert abs(flux / (3 * V) - 1) < 1e-7
    assert np.abs(r.volumes.sum(1) - g.volume).max() < 1e-15
    assert np.abs(r.edge_area.sum(1) - g.sqrtgram * 2 ** (g.D - 1)).max() < 1e-15


def test_torus_volume_and_area():
    Rr, rr = 0.25, 0.1
    r = XdgRules(CartesianGrid((12, 12, 12)), PolyLevelSet.torus([0.5, 0.5, 0.5], Rr, rr), n=8)
    assert abs(r.volumes[:, 0].sum() / (2 * np.pi ** 2 * Rr * rr ** 2) - 1) < 1e-7    # 2e-12 at 16^3
    assert abs(r.interface_area.sum() / (4 * np.pi ** 2 * Rr * rr) - 1) < 1e-5          # 4e-11 at 16^3
    assert (r.nspec[r.cut] == 2).all() and r.cut.sum() > 0
