"""Cartesian grids with the reference's numbering and affine metrics.

Mirrors Grid2D.Cartesian2DGrid / Grid3D.Cartesian3DGrid (Grid2D.cs:396ff,
Grid3D.cs:85-235: cell (i, j[, k]) has index (i*ny + j)[*nz + k], x slowest) and
the GridData metrics the operator reads (GridData_CellData.cs:293-387, 664-698;
GridData_EdgeData.cs:213-222, 2321, 2524-2580):
  detJ = prod(h_d / 2), invJ = diag(2 / h_d), Scaling = 1/sqrt(detJ),
  CellLengthScale = 2 |K| / (surface + boundary area)   (boundary faces count twice),
  SqrtGramian = face area / 2^(D-1), normals outward from the IN cell.
Edges: for each direction, all interior (and periodic) faces in cell order with
IN = lower cell, then the boundary faces (OUT = -1).  Local faces / trafo indices
follow Cube.Faces / Square.Faces: 0:-x 1:+x 2:+y 3:-y 4:+z 5:-z.
"""
import numpy as np

_FACE = {(0, 0): 0, (0, 1): 1, (1, 0): 3, (1, 1): 2, (2, 0): 5, (2, 1): 4}


class CartesianGrid:
    """Tensor-product grid: uniform (ncells, lo, hi) or from explicit node coordinates per
    direction (``nodes``, as Cartesian2DGrid/Cartesian3DGrid take them)."""

    def __init__(self, ncells=None, lo=None, hi=None, periodic=None, nodes=None):
        if nodes is None:
            D = len(ncells)
            lo = np.zeros(D) if lo is None else np.asarray(lo, float)
            hi = np.ones(D) if hi is None else np.asarray(hi, float)
            nodes = [np.linspace(lo[d], hi[d], int(ncells[d]) + 1) for d in range(D)]
        self.nodes = [np.asarray(x, float) for x in nodes]
        self.ncells = tuple(len(x) - 1 for x in self.nodes)
        D = self.D = len(self.ncells)
        assert D in (2, 3)
        self.lo = np.array([x[0] for x in self.nodes])
        self.hi = np.array([x[-1] for x in self.nodes])
        self.periodic = (False,) * D if periodic is None else tuple(bool(p) for p in periodic)
        n = np.array(self.ncells)
        self.J = int(np.prod(n))
        idx = np.indices(self.ncells).reshape(D, -1).T      # x slowest, last index fastest
        self.cell_lo = np.stack([self.nodes[d][idx[:, d]] for d in range(D)], 1)
        self.cell_hi = np.stack([self.nodes[d][idx[:, d] + 1] for d in range(D)], 1)
        hc = self.cell_hi - self.cell_lo                      # [J, D]
        self.hc = hc
        self.centers = 0.5 * (self.cell_lo + self.cell_hi)
        self.detJ = np.prod(hc / 2, axis=1)
        self.invJ = np.zeros((self.J, D, D))
        for d in range(D):
            self.invJ[:, d, d] = 2.0 / hc[:, d]
        self.scaling = 1.0 / np.sqrt(self.detJ)
        farea = np.stack([np.prod(np.delete(hc, d, axis=1), axis=1) for d in range(D)], 1)   # [J, D]
        strides = np.array([int(np.prod(n[d + 1:])) for d in range(D)])
        cells, trafo, normals, sqrtg = [], [], [], []
        bnd_area = np.zeros(self.J)
        allc = np.arange(self.J)
        for d in range(D):
            up = idx[:, d] + 1
            ok = up < n[d]
            nbr = allc + strides[d]
            if self.periodic[d]:
                wrap = ~ok
                nbr = np.where(wrap, allc - (n[d] - 1) * strides[d], nbr)
                ok = np.ones_like(ok)
            ci = allc[ok]; co = nbr[ok]
            cells.append(np.stack([ci, co], 1))
            trafo.append(np.tile([_FACE[(d, 1)], _FACE[(d, 0)]], (len(ci), 1)))
            nd = np.zeros((len(ci), D)); nd[:, d] = 1.0
            normals.append(nd)
            sqrtg.append(farea[ci, d] / 2 ** (D - 1))
        for d in range(D):
            if self.periodic[d]:
                continue
            for side in (0, 1):
                sel = allc[idx[:, d] == (n[d] - 1 if side else 0)]
                cells.append(np.stack([sel, -np.ones_like(sel)], 1))
                trafo.append(np.tile([_FACE[(d, side)], -1], (len(sel), 1)))
                nd = np.zeros((len(sel), D)); nd[:, d] = 1.0 if side else -1.0
                normals.append(nd)
                sqrtg.append(farea[sel, d] / 2 ** (D - 1))
                bnd_area[sel] += farea[sel, d]
        self.cells = np.ascontiguousarray(np.concatenate(cells).astype(np.int32))
        self.trafo = np.ascontiguousarray(np.concatenate(trafo).astype(np.int32))
        self.normals = np.ascontiguousarray(np.concatenate(normals))
        self.sqrtgram = np.ascontiguousarray(np.concatenate(sqrtg))
        self.E = len(self.cells)
        self.volume = np.prod(hc, axis=1)
        surf = 2 * farea.sum(axis=1)
        self.lenscale = 2.0 * self.volume / (surf + bnd_area)

    def cell_index(self, ijk):
        n = self.ncells
        j = 0
        for d in range(self.D):
            j = j * n[d] + ijk[d]
        return j
