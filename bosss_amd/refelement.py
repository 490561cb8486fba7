"""Reference elements: Gauss rules and the L2-orthonormal modal basis.

Host-side mirror of RefElement / Square / Cube / Basis (RefElement.cs:844-872,
1022-1040; Cube.cs:155-196; Square.cs:110-127; Basis.cs:83-93):

* volume rule: tensor product of the 1D Gauss rule with the fewest nodes whose
  order is >= the requested one; node index (i*NN + j)*NN + k has x = node k,
  y = node j, z = node i (Cube.cs:167-178); in 2D node i*NN + j has x = node j,
  y = node i (Square.cs:110-127);
* basis: all polynomials of total degree <= p, sorted by degree and then
  lexicographically in the exponents (e_x, e_y, e_z), each the product of
  normalised Legendre polynomials sqrt((2n+1)/2) P_n (the reference prints these
  coefficients in Cube_polynomials.cs / Square.cs);
* faces: 0:-x 1:+x 2:+y 3:-y 4:+z 5:-z (Cube.Faces, Square.Faces) with the face
  coordinate maps of TransformFaceCoordinates.

The nodes and weights come from numpy's Gauss-Legendre routine (Golub-Welsch
eigenvalue method with a Newton polish), the basis from the closed form above.
"""
import itertools

import numpy as np
from numpy.polynomial import legendre as npleg


def gauss_points_for_order(order):
    """Number of 1D Gauss points of the rule the reference picks for `order` (2K-1 >= order)."""
    return max(1, (order + 2) // 2)


def line_rule(order):
    K = gauss_points_for_order(order)
    x, w = npleg.leggauss(K)
    return x, w


def volume_rule(D, order):
    x, w = line_rule(order)
    K = len(x)
    if D == 1:
        return x[:, None].copy(), w.copy()
    if D == 2:
        nodes = np.zeros((K * K, 2)); wts = np.zeros(K * K)
        for i in range(K):
            for j in range(K):
                nodes[i * K + j] = (x[j], x[i]); wts[i * K + j] = w[i] * w[j]
        return nodes, wts
    nodes = np.zeros((K ** 3, 3)); wts = np.zeros(K ** 3)
    for i in range(K):
        for j in range(K):
            for k in range(K):
                idx = (i * K + j) * K + k
                nodes[idx] = (x[k], x[j], x[i]); wts[idx] = w[i] * w[j] * w[k]
    return nodes, wts


def modes(D, p):
    out = []
    for deg in range(p + 1):
        out += sorted(e for e in itertools.product(range(deg + 1), repeat=D) if sum(e) == deg)
    return out


def num_modes(D, p):
    return len(modes(D, p))


def _leg(n, x, deriv=False):
    c = np.zeros(n + 1); c[-1] = np.sqrt((2 * n + 1) / 2.0)
    if deriv:
        c = npleg.legder(c) if n > 0 else np.zeros(1)
    return npleg.legval(x, c)


def basis(D, p, x):
    """phi[k, n] and dphi[k, n, d] at reference points x[k, :]."""
    ms = modes(D, p)
    K = x.shape[0]
    phi = np.ones((K, len(ms)))
    dphi = np.ones((K, len(ms), D))
    for n, e in enumerate(ms):
        vals = [_leg(e[d], x[:, d]) for d in range(D)]
        ders = [_leg(e[d], x[:, d], True) for d in range(D)]
        phi[:, n] = np.prod(vals, axis=0)
        for d in range(D):
            t = np.ones(K)
            for dd in range(D):
                t = t * (ders[dd] if dd == d else vals[dd])
            dphi[:, n, d] = t
    return phi, dphi


def face_to_volume(D, face, s):
    """Map face reference coordinates s[k, D-1] to volume coordinates of `face`."""
    K = s.shape[0]
    xi = np.zeros((K, D))
    one = np.ones(K)
    if D == 2:
        table = {0: (-one, s[:, 0]), 1: (one, s[:, 0]), 2: (s[:, 0], one), 3: (s[:, 0], -one)}
    else:
        table = {0: (-one, s[:, 1], s[:, 0]), 1: (one, s[:, 1], s[:, 0]), 2: (s[:, 0], one, s[:, 1]),
                 3: (s[:, 0], -one, s[:, 1]), 4: (s[:, 0], s[:, 1], one), 5: (s[:, 0], s[:, 1], -one)}
    for d, col in enumerate(table[face]):
        xi[:, d] = col
    return xi


class Tables:
    """All reference-element tables the kernels need, for (D, p, quadrature order)."""

    def __init__(self, D, p, order):
        self.D, self.p, self.order = D, p, order
        self.vnodes, self.vw = volume_rule(D, order)
        self.fnodes, self.fw = volume_rule(D - 1, order)
        self.phi, self.dphi = basis(D, p, self.vnodes)
        self.N = self.phi.shape[1]
        self.Kv, self.Ke = len(self.vw), len(self.fw)
        self.ntrafo = 2 * D
        self.trace = np.zeros((2 * D, self.Ke, self.N))
        self.dtrace = np.zeros((2 * D, self.Ke, self.N, D))
        for f in range(2 * D):
            self.trace[f], self.dtrace[f] = basis(D, p, face_to_volume(D, f, self.fnodes))
