"""Pins of the oracle's reference-element tables (Gauss rules, orthonormal basis).

Reference tests restated: QuadratureTest/QuadratueAndProjectionTest.cs:62-105
(sum of weights, exactness), DerivativeTestMain.cs:888-941 (mass matrix = I).
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle.bosss_ref import RefModel, TABLES


def _rules():
    t = open(f"{TABLES}/line_quad_rules.txt").read().split()
    n = int(t[0]); i = 1; out = []
    for _ in range(n):
        order, K = int(t[i]), int(t[i + 1]); i += 2
        a = np.array(t[i:i + 2 * K], dtype=float).reshape(K, 2); i += 2 * K
        out.append((order, a[:, 0], a[:, 1]))
    return out


def test_line_rules_match_gauss_legendre():
    rules = _rules()
    assert len(rules) == 60
    # stored in the resource's own order, not sorted
    assert sorted(r[0] for r in rules) == list(range(1, 120, 2))
    for order, x, w in rules:
        K = len(x)
        assert order == 2 * K - 1
        xg, wg = npleg.leggauss(K)
        assert np.abs(x - xg).max() < 1e-15
        # the stored weights carry ~15 digits; the very high orders (>= 101) lose
        # accuracy in the stored data itself (up to 3.3e-11 at order 119)
        tol = 1e-14 if order < 100 else 1e-10
        assert np.abs(w - wg).max() < tol
        assert abs(w.sum() - 2.0) < 10 * tol


@pytest.mark.parametrize("order", [1, 5, 9, 21])
def test_line_rule_exactness(order):
    rule = [r for r in _rules() if r[0] == order][0]
    _, x, w = rule
    for deg in range(order + 1):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-13


@pytest.mark.parametrize("D,p", [(2, 2), (2, 4), (3, 2), (3, 3)])
def test_basis_is_normalised_legendre_product(D, p):
    m = RefModel(D, (2,) * D, p, 2 * p)
    rng = np.random.default_rng(D * 10 + p)
    x = rng.uniform(-1, 1, (50, D))
    phi, dphi = m.basis(x)
    # degree-sorted, lexicographic in (e_x, e_y, e_z)
    import itertools
    modes = []
    for deg in range(p + 1):
        modes += sorted(e for e in itertools.product(range(deg + 1), repeat=D) if sum(e) == deg)
    assert len(modes) == m.N
    for n, e in enumerate(modes):
        ref = np.ones(len(x))
        for d in range(D):
            c = np.zeros(e[d] + 1); c[-1] = np.sqrt((2 * e[d] + 1) / 2.0)
            ref *= npleg.legval(x[:, d], c)
        assert np.abs(phi[:, n] - ref).max() < 1e-13 * max(1, np.abs(ref).max())
        # derivative w.r.t. x_0 of the product
        c0 = np.zeros(e[0] + 1); c0[-1] = np.sqrt((2 * e[0] + 1) / 2.0)
        dref = npleg.legval(x[:, 0], npleg.legder(c0))
        for d in range(1, D):
            c = np.zeros(e[d] + 1); c[-1] = np.sqrt((2 * e[d] + 1) / 2.0)
            dref = dref * npleg.legval(x[:, d], c)
        assert np.abs(dphi[:, n, 0] - dref).max() < 1e-12 * max(1, np.abs(dref).max())


@pytest.mark.parametrize("D,p,order", [(2, 2, 4), (3, 2, 6), (3, 3, 6), (3, 3, 9)])
def test_mass_matrix_identity_and_rule_sizes(D, p, order):
    m = RefModel(D, (2,) * D, p, order)
    t = m.tables()
    M = np.einsum("kn,km,k->nm", t["phi"], t["phi"], t["vw"])
    assert np.abs(M - np.eye(m.N)).max() < 1e-13
    K1 = (order + 2) // 2  # Gauss rule with 2K-1 >= order
    assert m.Kv == K1 ** D and m.Ke == K1 ** (D - 1)
    # traces: face nodes on the face of the reference element
    for f in range(2 * D):
        Mf = np.einsum("kn,km,k->nm", t["trace"][f], t["trace"][f], t["fw"])
        assert np.allclose(Mf, Mf.T)


def test_config_quadrature_sizes():
    # SumOfMaxDegrees(2+2 -> 4), NonLinearWithoutParameters(2): 2p+p (QuadOrderFunc.cs)
    assert (RefModel(2, (2, 2), 2, 4).Kv, RefModel(2, (2, 2), 2, 4).Ke) == (9, 3)
    assert RefModel(3, (2, 2, 2), 3, 6).Kv == 64
    assert RefModel(3, (2, 2, 2), 2, 6).Kv == 64
    assert RefModel(3, (2, 2, 2), 3, 9).Kv == 125
    assert RefModel(3, (2, 2, 2), 3, 9).Ke == 25
