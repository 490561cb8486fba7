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
    def __init__(self, ncells, lo=None, hi=None, periodic=None):
        self.ncells = tuple(int(n) for n in ncells)
        D = self.D = len(self.ncells)
        assert D in (2, 3)
        self.lo = np.zeros(D) if lo is None else np.asarray(lo, float)
        self.hi = np.ones(D) if hi is None else np.asarray(hi, float)
        self.periodic = (False,) * D if periodic is None else tuple(bool(p) for p in periodic)
        n = np.array(self.ncells)
        self.J = int(np.prod(n))
        h = (self.hi - self.lo) / n
        self.h = h
        idx = np.indices(self.ncells).reshape(D, -1).T      # x slowest, last index fastest
        self.centers = self.lo + (idx + 0.5) * h
        dj = np.prod(h / 2)
        self.detJ = np.full(self.J, dj)
        self.invJ = np.zeros((self.J, D, D))
        for d in range(D):
            self.invJ[:, d, d] = 2.0 / h[d]
        self.scaling = np.full(self.J, 1.0 / np.sqrt(dj))
        farea = np.array([np.prod(np.delete(h, d)) for d in range(D)])
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
            sqrtg.append(np.full(len(ci), farea[d] / 2 ** (D - 1)))
        for d in range(D):
            if self.periodic[d]:
                continue
            for side in (0, 1):
                sel = allc[idx[:, d] == (n[d] - 1 if side else 0)]
                cells.append(np.stack([sel, -np.ones_like(sel)], 1))
                trafo.append(np.tile([_FACE[(d, side)], -1], (len(sel), 1)))
                nd = np.zeros((len(sel), D)); nd[:, d] = 1.0 if side else -1.0
                normals.append(nd)
                sqrtg.append(np.full(len(sel), farea[d] / 2 ** (D - 1)))
                bnd_area[sel] += farea[d]
        self.cells = np.ascontiguousarray(np.concatenate(cells).astype(np.int32))
        self.trafo = np.ascontiguousarray(np.concatenate(trafo).astype(np.int32))
        self.normals = np.ascontiguousarray(np.concatenate(normals))
        self.sqrtgram = np.ascontiguousarray(np.concatenate(sqrtg))
        self.E = len(self.cells)
        vol = np.prod(h)
        surf = 2 * farea.sum()
        self.lenscale = 1.0 / ((surf + bnd_area) * 0.5 / vol)

    def cell_index(self, ijk):
        n = self.ncells
        j = 0
        for d in range(self.D):
            j = j * n[d] + ijk[d]
        return j
