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
