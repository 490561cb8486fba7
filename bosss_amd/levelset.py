"""Polynomial level sets and Saye cut-cell quadrature for them (host side).

``PolyLevelSet`` is any polynomial level set psi(x) = sum_i C[i_0..i_{D-1}] x_0^i_0 ... x_{D-1}^i_{D-1}
in physical coordinates, of any degree per variable: rotated ellipsoids (cross terms), tori
(degree 4), products of spheres, or the projection of a DG level-set field on one cell.  The
reference represents its level set as a DG field (XdgPoisson3Control.cs:58-68) and generates
cut-cell rules with R. I. Saye's algorithm ("High-order quadrature methods for implicitly defined
surfaces and volumes in hyperrectangles", SIAM J. Sci. Comput. 37(2), A993-A1019, 2015;
Quadrature/SayeGaussRule_*.cs, ``CutCellQuadratureType = Saye``, AppControl.cs:1102).  The
functions here follow that algorithm for general polynomials:

  * constraints that keep one sign on a box are dropped (Bernstein bounds of the box and, when
    they straddle zero, of its 2^D and 4^D sub-boxes);
  * a height direction k is one in which every remaining constraint is monotone (the Bernstein
    bounds of d psi / d x_k exclude zero), chosen with the largest |d psi / d x_k| at the centre;
    the base constraints are the restrictions to the bottom and top faces, the base box is
    integrated recursively, and each vertical line is split at the (unique, monotone) root of
    every constraint, found by bisection to machine precision, with an n-point Gauss rule per piece;
  * a box with no height direction is split in half along its longest side (Saye's
    subdivision), to a fixed depth;
  * the interface rule takes the root on each vertical line with weight
    w_base |grad psi| / |d psi / d x_k| and normal grad psi / |grad psi|.

The separable quadratics of ``bosss_amd.xdg.QuadLevelSet`` are a special case; the closed-form
path there stays for them.
"""
import math

import numpy as np

MAX_DEPTH = 12


def _binom(n, k):
    return math.comb(n, k)


_BERN = {}


def _bern_tables(q):
    """binom(i, k), the exponents i - k and k, the mask k <= i, and M[k, m] = binom(m, k) /
    binom(q, k) (t^k in the degree-q Bernstein basis), for degree q."""
    if q not in _BERN:
        i = np.arange(q + 1)[:, None]
        k = np.arange(q + 1)[None, :]
        binom = np.array([[_binom(a, b) if b <= a else 0 for b in range(q + 1)] for a in range(q + 1)], float)
        M = np.array([[_binom(m, kk) / _binom(q, kk) if m >= kk else 0.0 for m in range(q + 1)] for kk in range(q + 1)])
        _BERN[q] = (binom, np.maximum(i - k, 0), np.broadcast_to(k, (q + 1, q + 1)), k <= i, M)
    return _BERN[q]


class PolyLevelSet:
    """psi(x) = sum C[i_0, ..., i_{D-1}] prod_d x_d^{i_d} (monomials in physical coordinates)."""

    def __init__(self, C):
        self.C = np.asarray(C, float)
        self.D = self.C.ndim

    # ---------------------------------------------------------------- constructors
    @classmethod
    def from_quadratic(cls, q):
        """The separable quadratic sum_d (a_d x_d^2 + b_d x_d) + c as a tensor polynomial."""
        D = q.D
        C = np.zeros((3,) * D)
        C[(0,) * D] = q.c
        for d in range(D):
            i1 = [0] * D; i1[d] = 1
            i2 = [0] * D; i2[d] = 2
            C[tuple(i1)] += q.b[d]
            C[tuple(i2)] += q.a[d]
        return cls(C)

    @classmethod
    def quadratic_form(cls, A, center, rhs=1.0):
        """(x - c)^T A (x - c) - rhs: with A symmetric positive definite an ellipsoid (any
        orientation: A = R diag(1 / a_d^2) R^T)."""
        A = np.asarray(A, float)
        c = np.asarray(center, float)
        D = len(c)
        C = np.zeros((3,) * D)
        C[(0,) * D] = c @ A @ c - rhs
        for d in range(D):
            e = [0] * D; e[d] = 1
            C[tuple(e)] += -2.0 * (A[d] @ c)
            for k in range(D):
                e2 = [0] * D; e2[d] += 1; e2[k] += 1
                C[tuple(e2)] += A[d, k]
        return cls(C)

    @classmethod
    def torus(cls, center, R, r):
        """Torus around the z axis through `center`, radii R > r:
        (|x - c|^2 + R^2 - r^2)^2 - 4 R^2 ((x - c_0)^2 + (y - c_1)^2) (degree 4 in each variable)."""
        c = np.asarray(center, float)
        s = PolyLevelSet.quadratic_form(np.eye(3), c, rhs=r * r - R * R)        # |x - c|^2 + R^2 - r^2
        cyl = PolyLevelSet.quadratic_form(np.diag([1.0, 1.0, 0.0]), c, rhs=0.0)  # (x-c0)^2 + (y-c1)^2
        return s * s + cyl * (-4.0 * R * R)

    def __mul__(self, other):
        if np.isscalar(other):
            return PolyLevelSet(self.C * other)
        D = self.D
        shape = tuple(a + b - 1 for a, b in zip(self.C.shape, other.C.shape))
        out = np.zeros(shape)
        for i in np.ndindex(self.C.shape):
            if self.C[i] == 0.0:
                continue
            sl = tuple(slice(i[d], i[d] + other.C.shape[d]) for d in range(D))
            out[sl] += self.C[i] * other.C
        return PolyLevelSet(out)

    def __add__(self, other):
        shape = tuple(max(a, b) for a, b in zip(self.C.shape, other.C.shape))
        out = np.zeros(shape)
        out[tuple(slice(0, s) for s in self.C.shape)] += self.C
        out[tuple(slice(0, s) for s in other.C.shape)] += other.C
        return PolyLevelSet(out)

    # ---------------------------------------------------------------- evaluation
    def __call__(self, X):
        X = np.atleast_2d(np.asarray(X, float))
        v = self.C
        # contract the last axis first: v[..., i] x_last^i
        for d in range(self.D - 1, -1, -1):
            p = X[:, d:d + 1] ** np.arange(v.shape[-1])[None, :]         # [K, q+1]
            if d == self.D - 1:
                v = np.einsum("...i,ki->k...", v, p)
            else:
                v = np.einsum("k...i,ki->k...", v, p)
        return v.reshape(-1)

    def derivative(self, d):
        q = self.C.shape[d]
        if q <= 1:
            return PolyLevelSet(np.zeros_like(self.C))
        sl = [slice(None)] * self.D
        sl[d] = slice(1, None)
        k = np.arange(1, q).reshape([-1 if e == d else 1 for e in range(self.D)])
        return PolyLevelSet(self.C[tuple(sl)] * k)

    def grad(self, X):
        return np.stack([self.derivative(d)(X) for d in range(self.D)], 1)

    def restrict(self, d, value):
        """psi with x_d = value: a polynomial in the other D - 1 variables."""
        p = value ** np.arange(self.C.shape[d])
        return PolyLevelSet(np.tensordot(self.C, p, axes=([d], [0])))

    def line_coeffs(self, k, Y):
        """For base points Y [L, D-1] (the coordinates other than k, in order), the coefficients
        [L, q_k + 1] of the univariate polynomial t -> psi(..., x_k = t, ...)."""
        Y = np.asarray(Y, float).reshape(-1, self.D - 1)
        v = np.moveaxis(self.C, k, -1)                               # other axes in order, then k
        if self.D == 1:
            return np.broadcast_to(v, (len(Y), len(v))).copy()
        for d in range(self.D - 2, -1, -1):
            p = Y[:, d:d + 1] ** np.arange(v.shape[-2])[None, :]
            v = np.einsum("...ij,li->l...j", v, p) if d == self.D - 2 else np.einsum("l...ij,li->l...j", v, p)
        return v

    # ---------------------------------------------------------------- bounds
    def bernstein(self, lo, hi):
        """Bernstein coefficients of psi on the box(es) [lo, hi] (lo, hi: [D] or [B, D]);
        their min and max bound psi there (convex hull property), exactly at the corners."""
        lo = np.atleast_2d(np.asarray(lo, float))
        hi = np.atleast_2d(np.asarray(hi, float))
        B = lo.shape[0]
        if B == 1:
            # one box: contract the leading axis and append the result; after D steps the axes are
            # back in order
            v = self.C
            for d in range(self.D):
                binom, ex_lo, ex_h, tri, M = _bern_tables(self.C.shape[d] - 1)
                T = np.where(tri, binom * lo[0, d] ** ex_lo * (hi[0, d] - lo[0, d]) ** ex_h, 0.0)
                v = np.tensordot(v, T @ M, axes=([0], [0]))
            return v[None]
        v = np.broadcast_to(self.C, (B,) + self.C.shape)
        for d in range(self.D):
            q = self.C.shape[d] - 1
            binom, ex_lo, ex_h, tri, M = _bern_tables(q)
            h = hi[:, d] - lo[:, d]
            # monomial x^i with x = lo + h t:  sum_{k <= i} binom(i, k) lo^(i-k) h^k t^k, then t^k in
            # the Bernstein basis of degree q (M)
            T = np.where(tri, binom * lo[:, d, None, None] ** ex_lo * h[:, None, None] ** ex_h, 0.0)
            TM = T @ M
            v = np.moveaxis(np.einsum("b...i,bim->b...m", np.moveaxis(v, d + 1, -1), TM), -1, d + 1)
        return v

    def bounds(self, lo, hi):
        b = self.bernstein(lo, hi).reshape(np.atleast_2d(lo).shape[0], -1)
        return b.min(1), b.max(1)

    def sign_on(self, lo, hi, levels=2):
        """+1 / -1 if psi keeps that strict sign on the box, 0 if it may change sign there: the
        Bernstein bounds of the box, then of its 2^D and 4^D equal sub-boxes (one vectorised call
        per level)."""
        lo = np.asarray(lo, float); hi = np.asarray(hi, float)
        for lev in range(levels + 1):
            m = 1 << lev
            idx = np.stack(np.meshgrid(*[np.arange(m)] * len(lo), indexing="ij"), -1).reshape(-1, len(lo))
            h = (hi - lo) / m
            mn, mx = self.bounds(lo + idx * h, lo + (idx + 1) * h)
            if (mn > 0).all():
                return 1
            if (mx < 0).all():
                return -1
            if (mn > 0).any() and (mx < 0).any():
                return 0                                  # both signs certainly present
        return 0

    def monotone(self, k, lo, hi):
        """d psi / d x_k keeps a strict sign on the box (Bernstein bounds)."""
        mn, mx = self.derivative(k).bounds(lo, hi)
        return mn[0] > 0 or mx[0] < 0


# -------------------------------------------------------------------- 1D pieces
def _gauss(n):
    return np.polynomial.legendre.leggauss(n)


def _poly1d_roots(c, lo, hi):
    """Real roots of sum c_i t^i strictly inside (lo, hi), sorted."""
    c = np.trim_zeros(np.asarray(c, float), "b")
    if len(c) <= 1:
        return np.zeros(0)
    r = np.roots(c[::-1])
    r = r[np.abs(r.imag) <= 1e-12 * max(1.0, np.abs(r).max())].real
    r = np.sort(r[(r > lo) & (r < hi)])
    # polish (double roots and clustered roots are left as found)
    for _ in range(3):
        f = np.polynomial.polynomial.polyval(r, c)
        df = np.polynomial.polynomial.polyval(r, np.polynomial.polynomial.polyder(c))
        step = np.where(df != 0, f / np.where(df != 0, df, 1.0), 0.0)
        r = np.where(np.abs(step) < (hi - lo), r - step, r)
    return np.sort(r[(r > lo) & (r < hi)])


def _monotone_roots(C, lo, hi):
    """Root in (lo, hi) of each line polynomial C[L, q+1] known to be monotone there (NaN if the
    end values have the same sign): bisection to machine precision, vectorised over lines."""
    P = np.polynomial.polynomial
    flo = P.polyval(np.full(len(C), lo), C.T, tensor=False)
    fhi = P.polyval(np.full(len(C), hi), C.T, tensor=False)
    has = (flo < 0) != (fhi < 0)
    has &= (flo != 0) & (fhi != 0)
    a = np.full(len(C), lo, float)
    b = np.full(len(C), hi, float)
    sa = np.sign(flo)
    for _ in range(200):
        m = 0.5 * (a + b)
        if np.all((m == a) | (m == b) | ~has):
            break
        fm = P.polyval(m, C.T, tensor=False)
        left = np.sign(fm) == sa
        a = np.where(left, m, a)
        b = np.where(left, b, m)
    r = 0.5 * (a + b)
    return np.where(has, r, np.nan)


def _gauss_pieces(breaks, n):
    """Gauss nodes and weights of the pieces between consecutive breaks [L, P+1] (zero-length
    pieces get zero weight).  Returns T [L, P n], W [L, P n]."""
    x, w = _gauss(n)
    t0, t1 = breaks[:, :-1], breaks[:, 1:]
    half = 0.5 * (t1 - t0)
    mid = 0.5 * (t1 + t0)
    T = (mid[:, :, None] + half[:, :, None] * x[None, None, :]).reshape(len(breaks), -1)
    W = (half[:, :, None] * w[None, None, :]).reshape(len(breaks), -1)
    return T, W


# ------------------------------------------------------------ Saye recursion
def _active(funcs, lo, hi):
    """The constraints that may change sign on the box."""
    return [f for f in funcs if f.sign_on(lo, hi) == 0]


def _split(lo, hi):
    d = int(np.argmax(hi - lo))
    mid = 0.5 * (lo[d] + hi[d])
    h1 = hi.copy(); h1[d] = mid
    l2 = lo.copy(); l2[d] = mid
    return (lo, h1), (l2, hi)


def box_nodes(funcs, lo, hi, n, depth=0):
    """Quadrature nodes X [K, m] and weights W [K] of the whole box [lo, hi], every 1D piece split
    where a constraint changes sign, so that each constraint keeps its sign on every piece."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    m = len(lo)
    funcs = _active(funcs, lo, hi)
    if m == 1:
        br = [lo[0], hi[0]]
        for f in funcs:
            br += list(_poly1d_roots(f.C, lo[0], hi[0]))
        br = np.sort(np.array(br))[None, :]
        T, W = _gauss_pieces(br, n)
        keep = W[0] > 0
        return T[0][keep][:, None], W[0][keep]
    if not funcs:
        return _tensor(lo, hi, n)
    cand = [k for k in range(m) if all(f.monotone(k, lo, hi) for f in funcs)]
    if not cand:
        if depth >= MAX_DEPTH:
            return _tensor(lo, hi, n)          # unresolvable sliver: plain rule (documented limit)
        (l1, h1), (l2, h2) = _split(lo, hi)
        X1, W1 = box_nodes(funcs, l1, h1, n, depth + 1)
        X2, W2 = box_nodes(funcs, l2, h2, n, depth + 1)
        return np.concatenate([X1, X2]), np.concatenate([W1, W2])
    ctr = 0.5 * (lo + hi)[None, :]
    k = max(cand, key=lambda i: min(abs(float(f.derivative(i)(ctr)[0])) for f in funcs))
    base = []
    for f in funcs:
        base.append(f.restrict(k, lo[k]))
        base.append(f.restrict(k, hi[k]))
    blo, bhi = np.delete(lo, k), np.delete(hi, k)
    Y, Wy = box_nodes(base, blo, bhi, n, depth)
    roots = [_monotone_roots(f.line_coeffs(k, Y), lo[k], hi[k]) for f in funcs]
    R = np.sort(np.stack(roots, 1), axis=1) if roots else np.zeros((len(Y), 0))
    R = np.where(np.isnan(R), hi[k], R)
    br = np.concatenate([np.full((len(Y), 1), lo[k]), R, np.full((len(Y), 1), hi[k])], 1)
    T, Wt = _gauss_pieces(br, n)
    P = T.shape[1]
    X = np.empty((len(Y) * P, m))
    X[:, np.arange(m) != k] = np.repeat(Y, P, axis=0)
    X[:, k] = T.reshape(-1)
    W = (Wy[:, None] * Wt).reshape(-1)
    keep = W > 0
    return X[keep], W[keep]


def _tensor(lo, hi, n):
    x, w = _gauss(n)
    m = len(lo)
    g = np.meshgrid(*[0.5 * (lo[i] + hi[i]) + 0.5 * (hi[i] - lo[i]) * x for i in range(m)], indexing="ij")
    wg = np.meshgrid(*[0.5 * (hi[i] - lo[i]) * w for i in range(m)], indexing="ij")
    return np.stack([a.reshape(-1) for a in g], 1), np.prod(np.stack([a.reshape(-1) for a in wg], 1), axis=1)


def surface_nodes(psi, lo, hi, n, depth=0):
    """Nodes X, surface weights W and unit normals grad psi / |grad psi| of {psi = 0} in the box."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    m = len(lo)
    empty = (np.zeros((0, m)), np.zeros(0), np.zeros((0, m)))
    if psi.sign_on(lo, hi) != 0:
        return empty
    cand = [k for k in range(m) if psi.monotone(k, lo, hi)]
    if not cand:
        if depth >= MAX_DEPTH:
            return empty
        (l1, h1), (l2, h2) = _split(lo, hi)
        parts = [surface_nodes(psi, l1, h1, n, depth + 1), surface_nodes(psi, l2, h2, n, depth + 1)]
        return tuple(np.concatenate([p[q] for p in parts]) for q in range(3))
    ctr = 0.5 * (lo + hi)[None, :]
    k = max(cand, key=lambda i: abs(float(psi.derivative(i)(ctr)[0])))
    if m == 1:
        r = _poly1d_roots(psi.C, lo[0], hi[0])
        X = r[:, None]
        G = psi.grad(X) if len(r) else np.zeros((0, 1))
        return X, np.ones(len(r)), G / np.abs(G) if len(r) else np.zeros((0, 1))
    base = [psi.restrict(k, lo[k]), psi.restrict(k, hi[k])]
    Y, Wy = box_nodes(base, np.delete(lo, k), np.delete(hi, k), n, depth)
    t = _monotone_roots(psi.line_coeffs(k, Y), lo[k], hi[k])
    ok = ~np.isnan(t)
    X = np.empty((ok.sum(), m))
    X[:, np.arange(m) != k] = Y[ok]
    X[:, k] = t[ok]
    G = psi.grad(X)
    gn = np.linalg.norm(G, axis=1)
    W = Wy[ok] * gn / np.abs(G[:, k])
    return X, W, G / gn[:, None]
