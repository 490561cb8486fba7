"""Batched uneven cell-block factorisation and block-Jacobi (host mirror of BlockJacobi).

BlockJacobi.InitImpl (BlockJacobi.cs:63-120) reads each cell's diagonal block
(BlockMsrMatrix.ReadBlock) and inverts it in place (IMatrixExtensions.InvertInPlace ->
LAPACK DGETRF + DGETRI); BlockJacobi.Solve (BlockJacobi.cs:148-197) sweeps
    x <- omega * invDiag (b - (A - D) x) + (1 - omega) x.
Here the factorisation of all blocks is one kernel launch on device-resident
blocks, whatever mix of block sizes the batch holds (e.g. 10- and 20-DOF XDG cells).
"""
import ctypes

import numpy as np

from ._lib import device_array, lib


class BlockBatch:
    """Device-resident batch of dense square blocks of uneven sizes (torch tensors)."""

    def __init__(self, sizes, device="cuda:0"):
        import torch
        self.sizes = np.asarray(sizes, dtype=np.int32)
        if self.sizes.min() < 1 or self.sizes.max() > 4096:
            raise ValueError("block sizes must be in 1..4096 (above 128: factored in global memory, no inverse)")
        n64 = self.sizes.astype(np.int64)
        off = np.zeros(len(n64), np.int64); off[1:] = np.cumsum(n64 * n64)[:-1]
        xoff = np.zeros(len(n64), np.int64); xoff[1:] = np.cumsum(n64)[:-1]
        self.nblk = len(self.sizes)
        self.nmax = int(self.sizes.max())
        self.total = int((n64 * n64).sum())
        self.ndof = int(n64.sum())
        self.dev = torch.device(device)
        self.n_t = torch.from_numpy(self.sizes).to(self.dev)
        self.off_t = torch.from_numpy(off).to(self.dev)
        self.xoff_t = torch.from_numpy(xoff).to(self.dev)
        self.off, self.xoff = off, xoff

    @staticmethod
    def _p(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)

    def _check(self, matrix=None, ipiv=None, info=None, vecs=()):
        """Sizes the kernels index with: packed blocks (total), pivots (ndof), info (nblk), vectors (ndof)."""
        import torch
        d = self.dev.index if self.dev.index is not None else torch.cuda.current_device()
        if matrix is not None:
            device_array(matrix, torch.float64, self.total, d, "blocks")
        if ipiv is not None:
            device_array(ipiv, torch.int32, self.ndof, d, "ipiv")
        if info is not None:
            device_array(info, torch.int32, self.nblk, d, "info")
        for t in vecs:
            device_array(t, torch.float64, self.ndof, d, "vector")

    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def factor(self, A, ipiv, info, inverse=False):
        """In-place LU (or inverse) of all blocks of A (float64 tensor, packed row-major); ipiv and
        info may be None (no pivots kept, no singularity report)."""
        self._check(A, ipiv, info)
        rc = lib().bosss_hip_block_lu_dev(None, self.nblk, self.nmax, self._p(self.n_t), self._p(self.off_t), self._p(A),
                                           self._p(ipiv), self._p(self.xoff_t), self._p(info), 1 if inverse else 0, self._stream())
        if rc != 0:
            raise RuntimeError("block LU failed (%d): %s" % (rc, lib().bosss_hip_block_last_error().decode()))

    def solve(self, LU, ipiv, x):
        if ipiv is None:
            raise ValueError("ipiv is required")
        self._check(LU, ipiv, None, (x,))
        rc = lib().bosss_hip_block_solve_dev(None, self.nblk, self._p(self.n_t), self._p(self.off_t), self._p(LU), self._p(ipiv),
                                              self._p(self.xoff_t), self._p(self.xoff_t), self._p(x), self._stream())
        if rc != 0:
            raise RuntimeError("block solve failed (%d): %s" % (rc, lib().bosss_hip_block_last_error().decode()))

    def gemv(self, Ainv, x, y, alpha=1.0, beta=0.0):
        self._check(Ainv, None, None, (x, y))
        rc = lib().bosss_hip_block_gemv_dev(None, self.nblk, self._p(self.n_t), self._p(self.off_t), self._p(Ainv), self._p(self.xoff_t),
                                             alpha, self._p(x), beta, self._p(y), self._stream())
        if rc != 0:
            raise RuntimeError("block gemv failed (%d): %s" % (rc, lib().bosss_hip_block_last_error().decode()))


class CellBlockPreconditioner:
    """Right preconditioner z = D^-1 v with D the diagonal cell blocks of shift I + dR/dU(U): the
    block-Jacobi of a Newton-Krylov step (BlockJacobi.InitImpl / Solve, BlockJacobi.cs:63-197, on the
    diagonal blocks of the linearised operator), without assembling the off-diagonal blocks.

    The blocks come from own-cell JVP probes of ``op`` (``bosss_hip_op_diag_blocks_dev``; an FP32
    operator probes in FP32) in transposed storage, which the probes write contiguously; inverting
    D^T in place in FP64 (batched LU pivots, Gauss-Jordan inverse) gives (D^-1)^T, the layout in
    which the apply (``bosss_hip_block_gemv_uniform_dev``) reads each block column by column,
    coalesced; it is kept in the apply precision ("fp32": rounded once to FP32)."""

    def __init__(self, op, U, shift, precision="fp32"):
        import time
        import torch
        t0 = time.perf_counter()
        J, Nt = op.grid.J, op.nvars * op.N
        self.op, self.J, self.Nt = op, J, Nt
        self.dtype = torch.float32 if precision == "fp32" else torch.float64
        B = torch.empty(J * Nt * Nt, dtype=torch.float64, device=U.device)
        op.diag_blocks_dev(U, B, alpha=1.0, shift=shift, transposed=True)       # D_j^T
        batch = BlockBatch([Nt] * J, device=U.device)
        info = torch.zeros(J, dtype=torch.int32, device=U.device)
        batch.factor(B, None, info, inverse=True)                               # (D_j^T)^-1 = (D_j^-1)^T
        bad = int((info != 0).sum())
        if bad:
            raise ArithmeticException("%d singular diagonal blocks" % bad)
        self.invT = B if self.dtype == torch.float64 else B.to(self.dtype)
        del B
        torch.cuda.synchronize(U.device)
        self.setup_seconds = time.perf_counter() - t0

    def __call__(self, v, z):
        import torch
        d = self.op.ctx.device
        device_array(v, self.dtype, self.J * self.Nt, d, "vector")
        device_array(z, self.dtype, self.J * self.Nt, d, "vector")
        prec = 0 if self.dtype == torch.float64 else 1
        rc = lib().bosss_hip_block_gemv_uniform_dev(self.op.ctx.h, self.J, self.Nt, prec, ctypes.c_void_p(self.invT.data_ptr()), 1.0,
                                                    ctypes.c_void_p(v.data_ptr()), 0.0, ctypes.c_void_p(z.data_ptr()),
                                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc != 0:
            raise RuntimeError("block gemv failed (%d): %s" % (rc, lib().bosss_hip_krylov_last_error().decode()))


class BlockJacobi:
    """Block-Jacobi smoother on an operator matrix in block-ELL layout (host mirror of BlockJacobi).

    Init (BlockJacobi.cs:63-120): copy the diagonal cell blocks, invert them all in one
    batched launch.  Solve (BlockJacobi.cs:148-197):
        x <- omega * invD (b - (A - D) x) + (1 - omega) x  ==  x + omega * invD (b - A x),
    computed as one block-ELL SpMV and one batched block GEMV per sweep.
    """

    def __init__(self, op, blocks, omega=1.0):
        import torch
        self.op = op
        self.omega = omega
        S, entries, _, _ = op.matrix_layout()
        Nt = op.nvars * op.N
        J = op.grid.J
        self.blocks = blocks
        self.batch = BlockBatch([Nt] * J, device=blocks.device)
        self.invD = blocks.view(J, S, Nt * Nt)[:, 0, :].contiguous().view(-1)
        self.info = torch.zeros(J, dtype=torch.int32, device=blocks.device)
        self.batch.factor(self.invD, None, self.info, inverse=True)
        self.r = torch.empty(op.ndof, dtype=torch.float64, device=blocks.device)
        self.z = torch.empty_like(self.r)

    def check(self):
        bad = int((self.info != 0).sum())
        if bad:
            raise ArithmeticException("%d singular diagonal blocks" % bad)

    def apply_inverse_diagonal(self, r, z):
        """z <- invD r (the block-Jacobi preconditioner)."""
        self.batch.gemv(self.invD, r, z)

    def Solve(self, x, b, sweeps=1):
        for _ in range(sweeps):
            self.r.copy_(b)
            self.op.spmv_dev(self.blocks, x, self.r, alpha=-1.0, beta=1.0)
            self.batch.gemv(self.invD, self.r, x, alpha=self.omega, beta=1.0)
        return x


class ArithmeticException(ArithmeticError):
    pass


def pcg(apply_A, b, precond, tol=1e-10, maxit=1000):
    """Preconditioned CG (torch, device resident); returns (x, iterations, relative residual)."""
    import torch
    x = torch.zeros_like(b)
    r = b.clone()
    z = torch.empty_like(b)
    precond(r, z)
    p = z.clone()
    rz = float(torch.dot(r, z))
    bn = float(torch.linalg.vector_norm(b))
    Ap = torch.empty_like(b)
    it = 0
    rel = 1.0
    while it < maxit:
        apply_A(p, Ap)
        alpha = rz / float(torch.dot(p, Ap))
        x.add_(p, alpha=alpha)
        r.add_(Ap, alpha=-alpha)
        it += 1
        rel = float(torch.linalg.vector_norm(r)) / bn
        if rel <= tol:
            break
        precond(r, z)
        rz_new = float(torch.dot(r, z))
        p.mul_(rz_new / rz).add_(z)
        rz = rz_new
    return x, it, rel


class Schwarz:
    """Additive Schwarz smoother over multi-cell blocks (host mirror of Schwarz,
    Schwarz.cs:985-1080; block factorisation as InitBlockSolver, Schwarz.cs:827-915).

    parts: lists of cell indices; with overlap > 0 each part grows by that many layers of
    neighbouring cells (the matrix pattern's neighbours).  Init reads every part's dense
    block from the operator matrix in block-ELL layout (one gather) and factors all parts in
    one batched launch (getrf semantics; parts above 128 DOF factor in global memory).
    Solve(x, b, sweeps): per sweep, Res = b - M x, then x += sum_parts R_p^T A_p^-1 R_p Res,
    the corrections of overlapping cells summed in part order (no overlap scaling: the
    reference's default EnableOverlapScaling = false).
    """

    def __init__(self, op, blocks, parts, overlap=0):
        import torch
        self.op = op
        self.blocks = blocks
        S, _, _, cols = op.matrix_layout()
        Nt = op.nvars * op.N
        J = op.grid.J
        parts = [sorted(set(int(c) for c in p)) for p in parts]
        for _ in range(int(overlap)):
            grown = []
            for p in parts:
                nb = set(p)
                for c in p:
                    nb.update(int(k) for k in cols[c] if k >= 0)
                grown.append(sorted(nb))
            parts = grown
        self.parts = parts
        sizes = [len(p) * Nt for p in parts]
        if max(sizes) > 4096:
            raise ValueError("a Schwarz block holds %d DOF; at most 4096" % max(sizes))
        # gather maps of every part's dense block into the block-ELL storage (-1: no coupling);
        # a column cell that occupies several slots of a row (a periodic direction of one or
        # two cells) contributes the sum of those blocks, one gather layer per repeat
        layers = []
        r = np.arange(Nt)
        for pi, p in enumerate(parts):
            m = len(p)
            slots = {}
            for ca in p:
                for s in range(S):
                    slots.setdefault((ca, int(cols[ca][s])), []).append(s)
            depth = max(len(v) for v in slots.values())
            while len(layers) < depth:
                layers.append([np.full(n * n, -1, np.int64) for n in sizes[:pi]])
            P = [np.full((m * Nt, m * Nt), -1, np.int64) for _ in range(len(layers))]
            for a, ca in enumerate(p):
                for b_, cb in enumerate(p):
                    for li, s in enumerate(slots.get((ca, cb), [])):
                        P[li][a * Nt:(a + 1) * Nt, b_ * Nt:(b_ + 1) * Nt] = (ca * S + s) * Nt * Nt + r[:, None] * Nt + r[None, :]
            for li in range(len(layers)):
                layers[li].append(P[li].ravel())
        self.batch = BlockBatch(sizes, device=blocks.device)
        flat = blocks.view(-1)
        vals = None
        for layer in layers:
            idx = torch.from_numpy(np.concatenate(layer)).to(blocks.device)
            v = flat[idx.clamp(min=0)] * (idx >= 0).to(blocks.dtype)
            vals = v if vals is None else vals + v
        self.LU = vals.contiguous()
        self.ipiv = torch.zeros(self.batch.ndof, dtype=torch.int32, device=blocks.device)
        self.info = torch.zeros(self.batch.nblk, dtype=torch.int32, device=blocks.device)
        self.batch.factor(self.LU, self.ipiv, self.info)
        # DOFs of the parts in part order, and for every DOF its positions there (fixed order)
        dof = np.concatenate([(np.asarray(p)[:, None] * Nt + np.arange(Nt)[None, :]).ravel() for p in parts])
        self.dof = torch.from_numpy(dof).to(blocks.device)
        order = np.argsort(dof, kind="stable")
        counts = np.bincount(dof, minlength=J * Nt)
        width = int(counts.max()) if len(counts) else 1
        back = np.full((J * Nt, width), -1, np.int64)
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        rank = np.arange(len(dof)) - np.repeat(start, counts)
        back[dof[order], rank] = order
        self.back = torch.from_numpy(back).to(blocks.device)
        self.res = torch.empty(op.ndof, dtype=torch.float64, device=blocks.device)

    def check(self):
        bad = int((self.info != 0).sum())
        if bad:
            raise ArithmeticException("%d singular Schwarz blocks" % bad)

    def apply(self, r, z):
        """z <- sum_parts R_p^T A_p^-1 R_p r (the additive Schwarz preconditioner)."""
        zp = r[self.dof].contiguous()
        self.batch.solve(self.LU, self.ipiv, zp)
        g = zp[self.back.clamp(min=0)] * (self.back >= 0).to(zp.dtype)
        z.copy_(g.sum(dim=1))

    def Solve(self, x, b, sweeps=1):
        import torch
        z = torch.empty_like(x)
        for _ in range(sweeps):
            self.res.copy_(b)
            self.op.spmv_dev(self.blocks, x, self.res, alpha=-1.0, beta=1.0)
            self.apply(self.res, z)
            x.add_(z)
        return x


def cartesian_parts(grid, agg):
    """Cells of a Cartesian grid grouped into boxes of agg[d] cells per direction (the last box
    of a direction takes the remainder)."""
    n = grid.ncells
    D = len(n)
    idx = np.indices(n).reshape(D, -1).T
    key = np.stack([idx[:, d] // agg[d] for d in range(D)], 1)
    nb = [(n[d] + agg[d] - 1) // agg[d] for d in range(D)]
    flat = np.ravel_multi_index(tuple(key.T), nb)
    return [np.nonzero(flat == k)[0] for k in range(int(np.prod(nb)))]
