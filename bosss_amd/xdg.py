"""Two-species XDG Poisson setup on Cartesian grids: level set, cut-cell quadrature,
species layout and cut-cell metrics.

Level sets are separable quadratics  psi(x) = sum_d (a_d x_d^2 + b_d x_d) + c  (spheres,
circles, planes).  Species A lives where psi < 0 and B where psi > 0 (the tracker
"-:A +:B", XdgPoisson3Main.cs:73).  The reference represents the level set as a DG field of
degree max(2, p) (XdgPoisson3Control.cs:58-68), so these quadratics are exact there too.

Cut-cell quadrature follows the dimension-reduction algorithm of R. I. Saye, "High-order
quadrature methods for implicitly defined surfaces and volumes in hyperrectangles",
SIAM J. Sci. Comput. 37(2), A993-A1019 (2015), which the reference's default
``CutCellQuadratureType = Saye`` implements (Quadrature/SayeGaussRule_*.cs,
AppControl.cs:1102): pick a height direction in which every constraint is monotone,
restrict the constraints to the bottom and top faces of the box to get the constraints of
the (D-1)-dimensional base, recurse, and integrate each vertical line with Gauss rules split
at the roots.  Boxes in which no direction is monotone are split at the vertex of the
offending quadratic.  Every 1D piece carries an n-point Gauss rule (n = 3 for the
quadrature order 2p = 4 of the XDG Poisson operator, XdgPoisson3Main.cs:164).

Metrics follow CutCellMetrics.ComputeNonAgglomeratedMetrics (CutCellMetrics.cs:150-360) and
the length-scale assembly of MultiphaseAgglomerator.LengthScaleAgg (MultiphaseAgglomerator.cs:
746-960) with agglomeration threshold 0 (no agglomeration pairs):
  LS_s(j) = Vol_s(j) / (InterfaceArea(j) + sum_{faces e of j} EdgeArea_s(e)),
  LS_s(j) = 1e10 * CellLengthScale(j) where Vol_s(j)/|K_j| <= 1e-10,
  LS_s(j) >= machine-eps * h_min(j).
"""
import numpy as np

from . import levelset
from .levelset import PolyLevelSet

SPECIES = ("A", "B")


class QuadLevelSet:
    """psi(x) = sum_d (a_d x_d^2 + b_d x_d) + c."""

    def __init__(self, a, b, c):
        self.a = np.asarray(a, float)
        self.b = np.asarray(b, float)
        self.c = float(c)
        self.D = len(self.a)

    @classmethod
    def sphere(cls, center, radius):
        ctr = np.asarray(center, float)
        return cls(np.ones(len(ctr)), -2 * ctr, float(ctr @ ctr) - radius ** 2)

    @classmethod
    def plane(cls, D, dim, offset):
        """psi = x_dim - offset."""
        b = np.zeros(D); b[dim] = 1.0
        return cls(np.zeros(D), b, -offset)

    def __call__(self, X):
        X = np.asarray(X, float)
        return (self.a * X * X + self.b * X).sum(-1) + self.c

    def grad(self, X):
        return 2 * self.a * np.asarray(X, float) + self.b

    def restrict(self, dim, value):
        return _restrict((self.a, self.b, self.c), dim, value)

    def as_tuple(self):
        return (self.a, self.b, self.c)


# ----------------------------------------------------------------- 1D helpers
def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def _restrict(f, dim, value):
    a, b, c = f
    return (np.delete(a, dim), np.delete(b, dim), c + a[dim] * value * value + b[dim] * value)


def _bounds(f, lo, hi):
    """Exact min/max of a separable quadratic over the box [lo, hi]."""
    a, b, c = f
    mn, mx = c, c
    for i in range(len(a)):
        cand = [a[i] * lo[i] ** 2 + b[i] * lo[i], a[i] * hi[i] ** 2 + b[i] * hi[i]]
        if a[i] != 0.0:
            v = -b[i] / (2 * a[i])
            if lo[i] < v < hi[i]:
                cand.append(a[i] * v * v + b[i] * v)
        mn += min(cand)
        mx += max(cand)
    return mn, mx


def _bounds_many(a, b, c, lo, hi):
    """_bounds for many boxes: a, b [m]; c scalar or [B]; lo, hi [B, m]."""
    vlo = a * lo * lo + b * lo
    vhi = a * hi * hi + b * hi
    mn = np.minimum(vlo, vhi)
    mx = np.maximum(vlo, vhi)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(a != 0, -b / np.where(a != 0, 2 * a, 1.0), np.nan)
    inside = (lo < v) & (v < hi)
    vv = a * v * v + b * v
    mn = np.where(inside, np.minimum(mn, vv), mn)
    mx = np.where(inside, np.maximum(mx, vv), mx)
    return mn.sum(-1) + c, mx.sum(-1) + c


def _monotone(f, i, lo, hi):
    a, b, _ = f
    if a[i] == 0.0:
        return True
    v = -b[i] / (2 * a[i])
    return not (lo < v < hi)


def _line_roots(a, b, c, lo, hi):
    """Roots of a t^2 + b t + c (c: array over lines) strictly inside (lo, hi); NaN where none."""
    c = np.asarray(c, float)
    r1 = np.full(c.shape, np.nan)
    r2 = np.full(c.shape, np.nan)
    if a == 0.0:
        if b != 0.0:
            r1 = -c / b
    else:
        disc = b * b - 4 * a * c
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0.0))
        # numerically stable pair of roots
        q = -0.5 * (b + np.copysign(sq, b if b != 0 else 1.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            r1 = np.where(ok, q / a, np.nan)
            r2 = np.where(ok & (q != 0), c / q, np.nan)
    out = []
    for r in (r1, r2):
        r = np.where((r > lo) & (r < hi), r, np.nan)
        out.append(r)
    return out


def _line_rule(funcs_1d, lo, hi, n):
    """Gauss rule along [lo, hi] for each line, split at the roots of every constraint.

    funcs_1d: list of (a, b, c[L]); returns T[L, P], W[L, P] (P = pieces * n; pieces of zero
    length carry zero weight)."""
    x, w = _gauss(n)
    L = len(np.atleast_1d(funcs_1d[0][2])) if funcs_1d else 1
    roots = []
    for a, b, c in funcs_1d:
        roots += _line_roots(a, b, np.broadcast_to(np.asarray(c, float), (L,)), lo, hi)
    if roots:
        R = np.sort(np.stack(roots, 1), axis=1)               # NaN sorts last
        R = np.where(np.isnan(R), hi, R)
    else:
        R = np.zeros((L, 0))
    br = np.concatenate([np.full((L, 1), lo), R, np.full((L, 1), hi)], 1)
    t0, t1 = br[:, :-1], br[:, 1:]
    half = 0.5 * (t1 - t0)
    mid = 0.5 * (t1 + t0)
    T = (mid[:, :, None] + half[:, :, None] * x[None, None, :]).reshape(L, -1)
    W = (half[:, :, None] * w[None, None, :]).reshape(L, -1)
    return T, W


# ------------------------------------------------------------ Saye recursion
def box_nodes(funcs, lo, hi, n):
    """Quadrature nodes of the box [lo, hi] (all of it), with every 1D piece split where a
    constraint in `funcs` changes sign, so that the sign of each constraint is constant
    along each piece.  Returns X[K, m], W[K]."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    m = len(lo)
    if m == 1:
        T, W = _line_rule([(f[0][0], f[1][0], f[2]) for f in funcs], lo[0], hi[0], n)
        keep = W[0] > 0
        return T[0][keep][:, None], W[0][keep]
    cand = [i for i in range(m) if all(_monotone(f, i, lo[i], hi[i]) for f in funcs)]
    if not cand:
        # split at the vertex of the first non-monotone constraint
        for f in funcs:
            for i in range(m):
                if not _monotone(f, i, lo[i], hi[i]):
                    v = -f[1][i] / (2 * f[0][i])
                    h1 = hi.copy(); h1[i] = v
                    l2 = lo.copy(); l2[i] = v
                    X1, W1 = box_nodes(funcs, lo, h1, n)
                    X2, W2 = box_nodes(funcs, l2, hi, n)
                    return np.concatenate([X1, X2]), np.concatenate([W1, W2])
    ctr = 0.5 * (lo + hi)

    def score(i):
        s = [abs(2 * f[0][i] * ctr[i] + f[1][i]) for f in funcs]
        return min(s) if s else 0.0
    k = max(cand, key=score)
    base = []
    for f in funcs:
        base.append(_restrict(f, k, lo[k]))
        base.append(_restrict(f, k, hi[k]))
    blo, bhi = np.delete(lo, k), np.delete(hi, k)
    Y, Wy = box_nodes(base, blo, bhi, n)
    line = []
    for a, b, c in funcs:
        cy = c + (np.delete(a, k) * Y * Y + np.delete(b, k) * Y).sum(1)
        line.append((a[k], b[k], cy))
    T, Wt = _line_rule(line, lo[k], hi[k], n)
    P = T.shape[1]
    X = np.empty((len(Y) * P, m))
    X[:, np.arange(m) != k] = np.repeat(Y, P, axis=0)
    X[:, k] = T.reshape(-1)
    W = (Wy[:, None] * Wt).reshape(-1)
    keep = W > 0
    return X[keep], W[keep]


def surface_nodes(psi, lo, hi, n):
    """Nodes, surface weights and unit normals (grad psi / |grad psi|) of {psi = 0} in the box."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    f = psi if isinstance(psi, tuple) else psi.as_tuple()
    m = len(lo)
    cand = [i for i in range(m) if _monotone(f, i, lo[i], hi[i])]
    if not cand:
        for i in range(m):
            if not _monotone(f, i, lo[i], hi[i]):
                v = -f[1][i] / (2 * f[0][i])
                h1 = hi.copy(); h1[i] = v
                l2 = lo.copy(); l2[i] = v
                parts = [surface_nodes(f, lo, h1, n), surface_nodes(f, l2, hi, n)]
                return tuple(np.concatenate([p[q] for p in parts]) for q in range(3))
    ctr = 0.5 * (lo + hi)
    k = max(cand, key=lambda i: abs(2 * f[0][i] * ctr[i] + f[1][i]))
    a, b, c = f
    if m == 1:
        r = [x for x in _line_roots(a[0], b[0], np.array([c]), lo[0], hi[0]) if not np.isnan(x[0])]
        X = np.array([[x[0]] for x in r]).reshape(-1, 1)
        G = 2 * a * X + b
        return X, np.ones(len(X)), G / np.abs(G)
    base = [_restrict(f, k, lo[k]), _restrict(f, k, hi[k])]
    Y, Wy = box_nodes(base, np.delete(lo, k), np.delete(hi, k), n)
    cy = c + (np.delete(a, k) * Y * Y + np.delete(b, k) * Y).sum(1)
    r1, r2 = _line_roots(a[k], b[k], cy, lo[k], hi[k])
    t = np.where(np.isnan(r1), r2, r1)
    ok = ~np.isnan(t)
    X = np.empty((ok.sum(), m))
    X[:, np.arange(m) != k] = Y[ok]
    X[:, k] = t[ok]
    G = 2 * a * X + b
    gn = np.linalg.norm(G, axis=1)
    W = Wy[ok] * gn / np.abs(G[:, k])
    return X, W, G / gn[:, None]


def tensor_nodes(lo, hi, n):
    x, w = _gauss(n)
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    m = len(lo)
    grids = np.meshgrid(*[0.5 * (lo[i] + hi[i]) + 0.5 * (hi[i] - lo[i]) * x for i in range(m)], indexing="ij")
    wg = np.meshgrid(*[0.5 * (hi[i] - lo[i]) * w for i in range(m)], indexing="ij")
    X = np.stack([g.reshape(-1) for g in grids], 1)
    W = np.prod(np.stack([g.reshape(-1) for g in wg], 1), axis=1)
    return X, W


# ------------------------------------------------------------- XDG problem
def _cell_bounds(psi, lo, hi):
    """Lower and upper bounds of psi on the boxes [lo, hi] (exact for QuadLevelSet)."""
    if isinstance(psi, PolyLevelSet):
        return psi.bounds(lo, hi)
    a, b, c = psi.as_tuple()
    return _bounds_many(a, b, c, lo, hi)


def _cell_rule(psi, lo, hi, n):
    if isinstance(psi, PolyLevelSet):
        return levelset.box_nodes([psi], lo, hi, n)
    return box_nodes([psi.as_tuple()], lo, hi, n)


def _face_rule(psi, d, x, tl, th, n):
    """Rule of the face x_d = x over the tangential box [tl, th], split at the interface, and the
    species (0: psi < 0, 1: psi > 0) of each node."""
    if isinstance(psi, PolyLevelSet):
        ff = psi.restrict(d, x)
        Y, W = levelset.box_nodes([ff], tl, th, n)
        return Y, W, (ff(Y) > 0).astype(int)
    ff = _restrict(psi.as_tuple(), d, x)
    Y, W = box_nodes([ff], tl, th, n)
    return Y, W, (((ff[0] * Y * Y + ff[1] * Y).sum(1) + ff[2]) > 0).astype(int)


def _interface_rule(psi, lo, hi, n):
    if isinstance(psi, PolyLevelSet):
        return levelset.surface_nodes(psi, lo, hi, n)
    return surface_nodes(psi.as_tuple(), lo, hi, n)


class XdgRules:
    """Species layout, quadrature items and metrics of a two-species XDG discretisation.

    Items (all node coordinates in the reference coordinates [-1,1]^D of the cell they
    belong to, weights physical):
      volume:    (cell, species, nodes, weights)
      face:      (IN cell, OUT cell or -1, species, nodes in IN, nodes in OUT, weights,
                  normal IN->OUT), over the grid's edge list (IN/OUT as there)
      interface: (cut cell, nodes, weights, normals from A to B)
    """

    def __init__(self, grid, psi, n=3):
        self.grid, self.psi, self.n = grid, psi, n
        D, J = grid.D, grid.J
        self.D, self.J = D, J
        lo, hi = grid.cell_lo, grid.cell_hi
        ctr, hc = grid.centers, grid.hc
        # ---- cell classification: bounds of psi on every cell (exact for the separable
        # quadratics, Bernstein bounds for polynomials), then the candidates' rules decide
        mn, mx = _cell_bounds(psi, lo, hi)
        cand = np.where((mn < 0) & (mx > 0))[0]

        def to_ref(X, j):
            return 2.0 * (X - ctr[j]) / hc[j]

        cut_items, single = [], {}
        for j in cand:
            X, W = _cell_rule(psi, lo[j], hi[j], n)
            s = (psi(X) > 0).astype(int)
            present = [sp for sp in (0, 1) if (s == sp).any()]
            if len(present) == 1:
                single[j] = present[0]                 # bounds straddled zero, psi does not
                continue
            for sp in present:
                sel = s == sp
                cut_items.append((j, sp, to_ref(X[sel], j), W[sel]))
        self.cut = np.zeros(J, bool)
        self.cut[[it[0] for it in cut_items]] = True
        noncand = ~((mn < 0) & (mx > 0))                 # one species: A where psi <= 0 there, else B
        has_A = (noncand & (mn < 0)) | self.cut
        has_B = (noncand & ~(mn < 0)) | self.cut
        for j, sp in single.items():
            (has_A if sp == 0 else has_B)[j] = True
        self.has = np.stack([has_A, has_B], 1)                      # [J, 2]
        self.nspec = self.has.sum(1).astype(np.int32)
        self.spec0 = np.where(has_A, 0, 1).astype(np.int32)        # species of the first block

        # ---- volume items
        xr, wr = tensor_nodes(-np.ones(D), np.ones(D), n)           # reference rule for uncut cells
        unc = np.where(~self.cut)[0]
        vol = np.zeros((J, 2))
        u_spec = np.where(has_A[unc], 0, 1)
        vol[unc, u_spec] = grid.volume[unc]
        self._uncut_cells = unc
        cut_cells = np.where(self.cut)[0]
        for j, sp, _, W in cut_items:
            vol[j, sp] = W.sum()
        # assemble: uncut (vectorised) then cut
        nu, K0 = len(unc), len(wr)
        self.vol_cell = np.concatenate([unc, [it[0] for it in cut_items]]).astype(np.int32)
        self.vol_spec = np.concatenate([u_spec, [it[1] for it in cut_items]]).astype(np.int32)
        cnt = np.concatenate([np.full(nu, K0), [len(it[3]) for it in cut_items]]).astype(np.int64)
        self.vol_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        self.vol_nodes = np.concatenate([np.tile(xr, (nu, 1))] + [it[2] for it in cut_items])
        self.vol_w = np.concatenate([(wr[None, :] * grid.detJ[unc, None]).reshape(-1)] + [it[3] for it in cut_items])
        self.volumes = vol

        # ---- face items over the grid's edge list
        E = grid.E
        ein, eout = grid.cells[:, 0].astype(np.int64), grid.cells[:, 1].astype(np.int64)
        edir = np.argmax(np.abs(grid.normals), axis=1)
        eplus = grid.normals[np.arange(E), edir] > 0
        exf = np.where(eplus, hi[ein, edir], lo[ein, edir])           # face coordinate
        # bounds of psi on each face: psi restricted to x_d = exf over the other dims
        flo, fhi = lo[ein].copy(), hi[ein].copy()
        flo[np.arange(E), edir] = exf
        fhi[np.arange(E), edir] = exf
        fmn, fmx = _cell_bounds(psi, flo, fhi)
        ecut = (fmn < 0) & (fmx > 0)
        espec = np.where(fmn < 0, 0, 1)
        edge_area = np.zeros((E, 2))
        # uncut faces: tensor rule in each direction's reference face
        items = []        # (edge, species, X[K, D] physical, W[K])
        yr, wr1 = tensor_nodes(-np.ones(D - 1), np.ones(D - 1), n)
        for d in range(D):
            sel = np.where(~ecut & (edir == d))[0]
            if len(sel) == 0:
                continue
            tang = np.arange(D) != d
            half = 0.5 * (hi[ein[sel]][:, tang] - lo[ein[sel]][:, tang])    # [S, D-1]
            mid = 0.5 * (hi[ein[sel]][:, tang] + lo[ein[sel]][:, tang])
            X = np.empty((len(sel), len(wr1), D))
            X[:, :, tang] = mid[:, None, :] + half[:, None, :] * yr[None, :, :]
            X[:, :, d] = exf[sel][:, None]
            W = wr1[None, :] * np.prod(half, axis=1)[:, None]
            for i, e in enumerate(sel):
                items.append((e, espec[e], X[i], W[i]))
        for e in np.where(ecut)[0]:
            d = edir[e]
            tl, th = np.delete(lo[ein[e]], d), np.delete(hi[ein[e]], d)
            Y, W, sv = _face_rule(psi, d, exf[e], tl, th, n)
            for sp in (0, 1):
                m_ = sv == sp
                if m_.any():
                    X = np.empty((m_.sum(), D))
                    X[:, np.arange(D) != d] = Y[m_]
                    X[:, d] = exf[e]
                    items.append((e, sp, X, W[m_]))
        items.sort(key=lambda it: (it[0], it[1]))
        fe = np.array([it[0] for it in items], np.int64)
        self.face_edge = fe.astype(np.int32)
        self.face_spec = np.array([it[1] for it in items], np.int32)
        self.face_cells = np.ascontiguousarray(grid.cells[fe].astype(np.int32))
        self.face_normal = np.ascontiguousarray(grid.normals[fe])
        cnt = np.array([len(it[3]) for it in items], np.int64)
        self.face_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        Xall = np.concatenate([it[2] for it in items])
        self.face_w = np.concatenate([it[3] for it in items])
        rep = np.repeat(np.arange(len(items)), cnt)
        jin, jout = self.face_cells[rep, 0], self.face_cells[rep, 1]
        self.face_x = Xall
        self.face_nodes_in = 2.0 * (Xall - ctr[jin]) / hc[jin]
        jo = np.where(jout >= 0, jout, jin)
        xo = 2.0 * (Xall - ctr[jo]) / hc[jo]
        dd = edir[fe][rep]
        pl = eplus[fe][rep]
        xo[np.arange(len(xo)), dd] = np.where(pl, -1.0, 1.0)      # OUT's own side (periodic faces too)
        xo[jout < 0] = 0.0
        self.face_nodes_out = xo
        np.add.at(edge_area, (fe, self.face_spec), np.add.reduceat(self.face_w, self.face_off[:-1]) if len(cnt) else 0)
        self.edge_area = edge_area

        # ---- interface items
        icell, inodes, iw, inorm = [], [], [], []
        iarea = np.zeros(J)
        for j in cut_cells:
            X, W, N = _interface_rule(psi, lo[j], hi[j], n)
            if len(W):
                icell.append(j); inodes.append(to_ref(X, j)); iw.append(W); inorm.append(N)
                iarea[j] = W.sum()
        self.ifc_cell = np.array(icell, np.int32)
        self.ifc_off = np.concatenate([[0], np.cumsum([len(w) for w in iw])]).astype(np.int64)
        self.ifc_nodes = np.concatenate(inodes) if inodes else np.zeros((0, D))
        self.ifc_w = np.concatenate(iw) if iw else np.zeros(0)
        self.ifc_normal = np.concatenate(inorm) if inorm else np.zeros((0, D))
        self.interface_area = iarea

        # ---- species length scales (agglomeration threshold 0)
        surf = np.zeros((J, 2)) + iarea[:, None]
        for sp in (0, 1):
            np.add.at(surf[:, sp], grid.cells[:, 0], edge_area[:, sp])
            bnd = grid.cells[:, 1] >= 0
            np.add.at(surf[:, sp], grid.cells[bnd, 1], edge_area[bnd, sp])
        with np.errstate(divide="ignore", invalid="ignore"):
            ls = vol / surf
        frac = vol / grid.volume[:, None]
        hmin = hc.min(axis=1)
        ls = np.where(frac <= 1e-10, 1e10 * grid.lenscale[:, None], np.maximum(ls, np.finfo(float).eps * hmin[:, None]))
        self.lenscale = np.ascontiguousarray(ls)
        self.volume_fraction = frac

    # ------------------------------------------------------------- layout
    def dof_layout(self, N):
        """Coefficient offsets: cell j holds nspec[j] blocks of N (species spec0[j] first)."""
        sizes = self.nspec.astype(np.int64) * N
        off = np.concatenate([[0], np.cumsum(sizes)])
        return off, sizes

    def species_offset(self, j, sp, N):
        """Offset of species sp's N coefficients inside cell j's block, or -1."""
        if not self.has[j, sp]:
            return -1
        return 0 if (self.spec0[j] == sp) else N

    def block_layout(self, N):
        """Cell-block matrix layout (block-ELL, uneven blocks).

        Returns a dict: S slots per block row (0 = diagonal, 1 + f = neighbour across local
        face f), cols[J, S] (column cell or -1), boff[J, S] (int64 offset of the row-major
        n_j x n_col block), nentries, dof offsets, and fslot[nface, 2] (the slot holding OUT
        in IN's row, and IN in OUT's row, for every face item)."""
        g = self.grid
        D, J = g.D, g.J
        S = 1 + 2 * D
        cols = -np.ones((J, S), np.int32)
        cols[:, 0] = np.arange(J)
        ci, co = g.cells[:, 0], g.cells[:, 1]
        fi, fo = g.trafo[:, 0], g.trafo[:, 1]
        inner = co >= 0
        cols[ci[inner], 1 + fi[inner]] = co[inner]
        cols[co[inner], 1 + fo[inner]] = ci[inner]
        nrow = self.nspec.astype(np.int64) * N
        ncol = np.where(cols >= 0, nrow[np.maximum(cols, 0)], 0)
        sizes = nrow[:, None] * ncol
        boff = np.concatenate([[0], np.cumsum(sizes.reshape(-1))])[:-1].reshape(J, S).astype(np.int64)
        dof = np.concatenate([[0], np.cumsum(nrow)]).astype(np.int64)
        e = self.face_edge
        fslot = np.stack([1 + g.trafo[e, 0], np.where(g.cells[e, 1] >= 0, 1 + g.trafo[e, 1], -1)], 1).astype(np.int32)
        return {"S": S, "cols": np.ascontiguousarray(cols), "boff": np.ascontiguousarray(boff),
                "nentries": int(sizes.sum()), "dof": dof, "ndof": int(dof[-1]), "fslot": np.ascontiguousarray(fslot),
                "nrow": nrow.astype(np.int32)}

    def boundary_values(self, g_D):
        """Dirichlet data g_D(x) at every face node (zero on inner faces)."""
        gd = np.zeros(len(self.face_w))
        rep = np.repeat(np.arange(len(self.face_spec)), np.diff(self.face_off))
        bnd = self.face_cells[rep, 1] < 0
        if g_D is not None and bnd.any():
            gd[bnd] = g_D(self.face_x[bnd])
        return gd


def desc_arrays(rules, p, g_D=None):
    """The arrays of the C descriptor ``bosss_xdg_desc`` for these cut-cell rules at degree p,
    keyed by short names (contiguous, with the descriptor's element types; ``gd`` is None
    without Dirichlet data)."""
    from .refelement import modes
    ex = np.ascontiguousarray(np.array(modes(rules.D, p), np.int32))
    lay = rules.block_layout(len(ex))
    g = rules.grid
    gd = rules.boundary_values(g_D) if g_D is not None else None
    c = np.ascontiguousarray
    return dict(ex=ex, nspec=c(rules.nspec, np.int32), spec0=c(rules.spec0, np.int32), invJ=c(g.invJ, np.float64),
                scal=c(g.scaling, np.float64), ls=c(rules.lenscale, np.float64), cols=c(lay["cols"], np.int32),
                boff=c(lay["boff"], np.int64), vc=c(rules.vol_cell, np.int32), vs=c(rules.vol_spec, np.int32),
                vo=c(rules.vol_off, np.int64), vn=c(rules.vol_nodes, np.float64), vw=c(rules.vol_w, np.float64),
                fc=c(rules.face_cells, np.int32), fsl=c(lay["fslot"], np.int32), fsp=c(rules.face_spec, np.int32),
                fo=c(rules.face_off, np.int64), fni=c(rules.face_nodes_in, np.float64), fno=c(rules.face_nodes_out, np.float64),
                fw=c(rules.face_w, np.float64), fnm=c(rules.face_normal, np.float64),
                gd=None if gd is None else c(gd, np.float64),
                ic=c(rules.ifc_cell, np.int32), io=c(rules.ifc_off, np.int64), inn=c(rules.ifc_nodes, np.float64),
                iw=c(rules.ifc_w, np.float64), inm=c(rules.ifc_normal, np.float64))


class XdgOperator:
    """Two-species XDG Poisson operator on the GPU (C ABI ``bosss_hip_xdg_*``).

    Mirrors what XdgPoisson3 builds (XdgPoisson3Main.cs:148-178): XLaplace_Bulk +
    XLaplace_Interface (mode SIP) + RHSSource with per-species constant sources, Dirichlet
    data g_D on the boundary, agglomeration threshold 0.  ``mu`` follows the reference's
    sign convention (MU_A, MU_B < 0 for the usual Poisson problem)."""

    def __init__(self, rules, p, mu, source=(0.0, 0.0), penalty=2.0, g_D=None, device=0):
        import ctypes
        from ._lib import lib, XdgDesc
        from .operator import Context
        self.rules, self.p = rules, p
        self.ctx = Context.get(device)
        D, J = rules.D, rules.J
        keep = desc_arrays(rules, p, g_D)
        ex = keep["ex"]
        N = self.N = len(ex)
        lay = self.layout = rules.block_layout(N)
        self._keep = keep

        def P(a, t):
            return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))
        d_, i_, l_ = ctypes.c_double, ctypes.c_int32, ctypes.c_int64
        desc = XdgDesc(D=D, degree=p, N=N, J=J, mode_exponents=P(ex, i_), nspec=P(keep["nspec"], i_), spec0=P(keep["spec0"], i_),
                       invJ=P(keep["invJ"], d_), scaling=P(keep["scal"], d_), lscale=P(keep["ls"], d_), penalty_base=float(penalty),
                       mu=(ctypes.c_double * 2)(*mu), source=(ctypes.c_double * 2)(*source), S=lay["S"],
                       cols=P(keep["cols"], i_), boff=P(keep["boff"], l_),
                       nvol=len(keep["vc"]), vol_cell=P(keep["vc"], i_), vol_spec=P(keep["vs"], i_), vol_off=P(keep["vo"], l_),
                       vol_nodes=P(keep["vn"], d_), vol_w=P(keep["vw"], d_),
                       nface=len(keep["fsp"]), face_cells=P(keep["fc"], i_), face_slot=P(keep["fsl"], i_), face_spec=P(keep["fsp"], i_),
                       face_off=P(keep["fo"], l_), face_nodes_in=P(keep["fni"], d_), face_nodes_out=P(keep["fno"], d_),
                       face_w=P(keep["fw"], d_), face_normal=P(keep["fnm"], d_), face_gD=P(keep["gd"], d_),
                       nifc=len(keep["ic"]), ifc_cell=P(keep["ic"], i_), ifc_off=P(keep["io"], l_), ifc_nodes=P(keep["inn"], d_),
                       ifc_w=P(keep["iw"], d_), ifc_normal=P(keep["inm"], d_))
        h = ctypes.c_void_p()
        self._check(lib().bosss_hip_xdg_create(self.ctx.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        nd, ne = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().bosss_hip_xdg_info(h, ctypes.byref(nd), ctypes.byref(ne)))
        self.ndof, self.nentries = nd.value, ne.value

    @staticmethod
    def _check(rc):
        from ._lib import lib, BosssError
        if rc != 0:
            raise BosssError(rc, lib().bosss_hip_xdg_last_error().decode())
        return rc

    def close(self):
        from ._lib import lib
        if getattr(self, "h", None):
            lib().bosss_hip_xdg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream(stream):
        import torch
        return (stream or torch.cuda.current_stream()).cuda_stream

    def new_matrix(self):
        import torch
        return torch.zeros(self.nentries, dtype=torch.float64, device="cuda:%d" % self.ctx.device)

    def _check_args(self, blocks, *vecs):
        """Matrix (nentries or more FP64 entries) and vectors (ndof FP64 entries) on this device."""
        import torch
        from ._lib import device_array
        device_array(blocks, torch.float64, self.nentries, self.ctx.device, "blocks", at_least=True)
        for t in vecs:
            device_array(t, torch.float64, self.ndof, self.ctx.device, "vector")

    def assemble_dev(self, blocks, affine, stream=None):
        """blocks <- M, affine <- b with R(u) = M u + b."""
        from ._lib import lib
        self._check_args(blocks, affine)
        self._check(lib().bosss_hip_xdg_assemble_dev(self.h, blocks.data_ptr(), affine.data_ptr(), self._stream(stream)))

    def spmv_dev(self, blocks, x, y, alpha=1.0, beta=0.0, stream=None):
        from ._lib import lib
        self._check_args(blocks, x, y)
        self._check(lib().bosss_hip_xdg_spmv_dev(self.h, blocks.data_ptr(), alpha, x.data_ptr(), beta, y.data_ptr(),
                                                 self._stream(stream)))

    def solve_dev(self, blocks, b, x, tol=1e-10, maxit=5000, precision="fp64", use_graph=True, check_every=16, stream=None):
        """Block-Jacobi PCG for M x = b (x: initial guess, overwritten).  Returns (iterations, outer, relres)."""
        import ctypes
        from ._lib import lib, BOSSS_FP64, BOSSS_FP32
        it, outer, rel = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        prec = BOSSS_FP64 if precision == "fp64" else BOSSS_FP32
        self._check_args(blocks, b, x)
        self._check(lib().bosss_hip_xdg_solve_dev(self.h, blocks.data_ptr(), b.data_ptr(), x.data_ptr(), tol, maxit, prec,
                                                  int(bool(use_graph)), check_every, ctypes.byref(it), ctypes.byref(outer),
                                                  ctypes.byref(rel), self._stream(stream)))
        return it.value, outer.value, rel.value

    def set_tile(self, t, precision="fp64"):
        """Fix the solver's tile shape (1, 2 or 4 groups of four cell rows per workgroup) or ask for
        a new tuning at the next solve (0)."""
        from ._lib import lib, BOSSS_FP64, BOSSS_FP32
        self._check(lib().bosss_hip_xdg_set_tile(self.h, BOSSS_FP64 if precision == "fp64" else BOSSS_FP32, int(t)))

    def tiles(self, precision="fp64"):
        """(chosen T or 0, {T: timed ms per apply + update}) of the solver's tile autotuner."""
        import ctypes
        from ._lib import lib, BOSSS_FP64, BOSSS_FP32
        c, ms = ctypes.c_int32(), (ctypes.c_float * 3)()
        self._check(lib().bosss_hip_xdg_tiles(self.h, BOSSS_FP64 if precision == "fp64" else BOSSS_FP32, ctypes.byref(c), ms))
        return c.value, {t: round(float(m), 5) for t, m in zip((1, 2, 4), ms) if m >= 0}
