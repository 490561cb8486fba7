"""GPU batched uneven block LU / inverse / solve against the oracle's getf2 restatement
and against LAPACK (scipy.linalg.lu_factor) on the host."""
import sys

import numpy as np
import pytest

from oracle.bosss_ref import ref_block_lu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_blocks(sizes, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, (n, n)) + n * 0.2 * np.eye(n) * rng.uniform(0.5, 1.5) for n in sizes]


@pytest.mark.parametrize("sizes", [[10] * 50 + [20] * 7, [1, 2, 3, 5, 8, 13, 21, 34, 55, 64, 65, 100], [50] * 40])
def test_block_lu_matches_oracle_pivots_and_factors(sizes):
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case block_lu", len(sizes), "blocks, max", max(sizes))
    blocks = make_blocks(sizes, 1)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, off, poff = ref_block_lu(sizes, blocks)
    LU = A.cpu().numpy()
    assert (info.cpu().numpy() == inforef).all()
    assert (ipiv.cpu().numpy() == pivref + 1).all()          # LAPACK 1-based ipiv
    assert np.abs(LU - LUref).max() < 1e-12 * np.abs(LUref).max()
    # solve against numpy
    x = np.random.default_rng(2).uniform(-1, 1, bb.ndof)
    xt = torch.from_numpy(x.copy()).to(bb.dev)
    bb.solve(A, ipiv, xt)
    xs = xt.cpu().numpy()
    for b, n in enumerate(sizes):
        o = bb.xoff[b]
        ref = np.linalg.solve(blocks[b], x[o:o + n])
        assert np.abs(xs[o:o + n] - ref).max() < 1e-11 * max(1, np.abs(ref).max())


def test_block_inverse_and_gemv():
    import torch
    from bosss_amd.blocks import BlockBatch
    sizes = [10] * 31 + [20] * 5 + [100] * 2
    say("case block_inverse", len(sizes))
    blocks = make_blocks(sizes, 3)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, None, info, inverse=True)
    x = np.random.default_rng(4).uniform(-1, 1, bb.ndof)
    xt = torch.from_numpy(x).to(bb.dev)
    y = torch.zeros_like(xt)
    bb.gemv(A, xt, y)
    torch.cuda.synchronize()
    inv = A.cpu().numpy()
    ys = y.cpu().numpy()
    for b, n in enumerate(sizes):
        Ainv = inv[bb.off[b]:bb.off[b] + n * n].reshape(n, n)
        assert np.abs(Ainv @ blocks[b] - np.eye(n)).max() < 1e-12 * n
        o = bb.xoff[b]
        assert np.abs(ys[o:o + n] - np.linalg.solve(blocks[b], x[o:o + n])).max() < 1e-11
    assert (info.cpu().numpy() == 0).all()


def test_singular_block_reports_info():
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case singular block")
    sizes = [4, 4, 4]
    blocks = make_blocks(sizes, 5)
    blocks[1][:, 2] = 0.0       # zero column -> zero pivot at j = 2 (INFO = 3)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    _, _, inforef, _, _ = ref_block_lu(sizes, blocks)
    assert list(info.cpu().numpy()) == [0, 3, 0] == list(inforef)


@pytest.mark.parametrize("sizes", [[129, 200, 10, 257, 20], [512, 300], [700, 64, 515]])
def test_large_blocks_factor_in_global_memory(sizes):
    """Blocks above 128 DOF (multi-cell Schwarz blocks): the global-memory getf2 gives the
    oracle's pivots and factors, solves match numpy, small blocks in the same batch too."""
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case large blocks", sizes)
    blocks = make_blocks(sizes, 7)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, off, poff = ref_block_lu(sizes, blocks)
    assert (info.cpu().numpy() == inforef).all() and (inforef == 0).all()
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    LU = A.cpu().numpy()
    assert np.abs(LU - LUref).max() < 1e-12 * np.abs(LUref).max()
    x = np.random.default_rng(3).uniform(-1, 1, bb.ndof)
    xt = torch.from_numpy(x.copy()).to(bb.dev)
    bb.solve(A, ipiv, xt)
    xs = xt.cpu().numpy()
    for b, n in enumerate(sizes):
        o = bb.xoff[b]
        ref = np.linalg.solve(blocks[b], x[o:o + n])
        assert np.abs(xs[o:o + n] - ref).max() < 1e-10 * max(1, np.abs(ref).max())


@pytest.mark.parametrize("col", [36, 37])                      # first or second column of a pair
@pytest.mark.parametrize("sizes", [[150, 140], [600, 520]])   # one-workgroup form; two launches per column
def test_large_block_zero_pivot_and_inverse_limit(sizes, col):
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case large block singular", sizes, "column", col)
    blocks = make_blocks(sizes, 9)
    blocks[0][:, col] = 0.0                 # zero column -> zero pivot at j = col (INFO = col + 1)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, _, _ = ref_block_lu(sizes, blocks)
    assert list(info.cpu().numpy()) == [col + 1, 0] == list(inforef)
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    # getf2 skips the zero column and goes on: the factors after it match too
    assert np.abs(A.cpu().numpy() - LUref).max() < 1e-12 * np.abs(LUref).max()
    with pytest.raises(RuntimeError):
        bb.factor(A, None, info, inverse=True)   # inverses are limited to 128 DOF


def zero_diagonal_block(n, j, seed):
    """A block whose diagonal entry is exactly zero when getf2 reaches column j while a row below
    holds a non-zero pivot: columns < j are unit vectors (no pivoting, no update of rows >= j),
    a[j][j] = 0, the rest random with a dominant diagonal."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (n, n)) + n * 0.2 * np.eye(n)
    A[:, :j] = 0.0
    A[:j, :j] = np.eye(j)
    A[j, j] = 0.0
    return A


@pytest.mark.parametrize("col", [36, 37])
@pytest.mark.parametrize("sizes", [[150, 140], [600, 520]])   # one-workgroup form; two launches per column
def test_large_block_zero_diagonal_swaps_in_the_pivot(sizes, col):
    """The row swap overwrites the pivot entry: every thread of the workgroup must have taken the
    same non-zero decision before it.  Factors and pivots equal the oracle's getf2."""
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case large block zero diagonal", sizes, "column", col)
    blocks = [zero_diagonal_block(n, col, 11 + i) for i, n in enumerate(sizes)]
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, _, poff = ref_block_lu(sizes, blocks)
    assert (inforef == 0).all() and (info.cpu().numpy() == 0).all()
    assert all(pivref[poff[b] + col] != col for b in range(len(sizes)))   # a real swap at column col
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    assert np.abs(A.cpu().numpy() - LUref).max() < 1e-12 * np.abs(LUref).max()


@pytest.mark.parametrize("sizes", [list(range(1, 17)) * 3, [10] * 61 + [20] * 9, [20, 19, 3, 17, 7, 12] * 5])
def test_register_kernel_small_blocks(sizes):
    """Blocks up to 20 DOF factor with their rows in registers, 64 / G blocks per wavefront (batches
    whose largest block is <= 16 and <= 20): pivots and factors equal the oracle's getf2,
    zero pivots (a zero column, even and odd) set INFO and the factorisation continues, a zero
    diagonal entry with a non-zero pivot below is swapped in, exactly tied candidates go to the
    first row, a NaN is never chosen as pivot; the inverses satisfy A inv(A) = I."""
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case register-kernel blocks", len(sizes), "blocks, max", max(sizes))
    blocks = make_blocks(sizes, 21)
    big = [b for b, n in enumerate(sizes) if n >= 6]
    blocks[big[0]][:, 2] = 0.0                       # zero pivot at an even column
    blocks[big[1]][:, 5] = 0.0                       # and at an odd one
    blocks[big[2]] = zero_diagonal_block(sizes[big[2]], 3, 5)
    blocks[big[3]][4, 0] = np.nan                    # NaN below the diagonal in column 0
    nt = sizes[big[4]]                               # exactly tied pivot candidates in column 0
    blocks[big[4]][:, 0] = np.where(np.arange(nt) % 3 == 1, 4.0, -4.0) * (np.arange(nt) % 2 * 2 - 1)
    bb = BlockBatch(sizes)
    A0 = np.concatenate([b.ravel() for b in blocks])
    A = torch.from_numpy(A0.copy()).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, off, poff = ref_block_lu(sizes, blocks)
    assert (info.cpu().numpy() == inforef).all()
    assert inforef[big[0]] == 3 and inforef[big[1]] == 6 and inforef[big[2]] == 0 and pivref[poff[big[4]]] == 0
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    LU = A.cpu().numpy()
    fin = np.isfinite(LUref)
    assert (np.isfinite(LU) == fin).all()
    assert np.abs(LU[fin] - LUref[fin]).max() < 1e-12 * np.abs(LUref[fin]).max()
    # inverses of the regular blocks
    ok = [b for b in range(len(sizes)) if inforef[b] == 0 and b != big[3]]
    Ai = torch.from_numpy(np.concatenate([blocks[b].ravel() for b in ok])).to(bb.dev)
    bi = BlockBatch([sizes[b] for b in ok])
    infoi = torch.zeros(bi.nblk, dtype=torch.int32, device=bi.dev)
    bi.factor(Ai, None, infoi, inverse=True)
    torch.cuda.synchronize()
    inv = Ai.cpu().numpy()
    assert (infoi.cpu().numpy() == 0).all()
    for k, b in enumerate(ok):
        n = sizes[b]
        Ainv = inv[bi.off[k]:bi.off[k] + n * n].reshape(n, n)
        assert np.abs(blocks[b] @ Ainv - np.eye(n)).max() < 1e-12 * n * max(1.0, np.abs(Ainv).max())


@pytest.mark.parametrize("sizes", [[9, 12], [40, 33], [150, 140], [600, 520]])   # registers, LDS, one workgroup, split
def test_nan_row_follows_idamax(sizes):
    """A NaN below the diagonal in column 0 turns that row into NaNs; when the elimination reaches
    the row's own column its NaN diagonal entry is the pivot (IDAMAX never moves past a NaN first
    element) and INFO stays 0 (NaN is not zero): the oracle's pivots in every kernel form."""
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case NaN pivots", sizes)
    blocks = make_blocks(sizes, 31)
    blocks[0][4, 0] = np.nan
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, off, poff = ref_block_lu(sizes, blocks)
    assert (info.cpu().numpy() == inforef).all() and (inforef == 0).all()
    assert pivref[4] == 4                                    # the NaN row is its column's pivot
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    LU = A.cpu().numpy()
    n1 = sizes[1]
    b1 = slice(off[1], off[1] + n1 * n1)                    # the second block is untouched by the NaN
    assert np.abs(LU[b1] - LUref[b1]).max() < 1e-12 * np.abs(LUref[b1]).max()


@pytest.mark.parametrize("sizes", [[21, 32, 33, 64, 65, 100, 112, 113, 128], [100] * 300])
def test_workgroup_tile_blocks(sizes):
    """Blocks of 21..128 DOF (one workgroup per block, the block in registers, row exchanges as
    bookkeeping): pivots and INFO equal the oracle's getf2 and the factors agree to rounding, with a
    zero column, a zero diagonal entry with a pivot below, exactly tied pivot candidates in column 0
    and a NaN below the diagonal; the inverses (in-place Gauss-Jordan) satisfy A inv(A) = I; a
    singular block asked for its inverse keeps getf2's factors and INFO instead."""
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case workgroup-tile blocks", len(sizes), "blocks, max", max(sizes))
    blocks = make_blocks(sizes, 41)
    blocks[0][:, 7] = 0.0                                  # zero pivot at column 7
    blocks[1] = zero_diagonal_block(sizes[1], 5, 3)        # zero diagonal, pivot swapped in
    blocks[2][:, 0] = np.where(np.arange(sizes[2]) % 3 == 1, 4.0, -4.0) * (np.arange(sizes[2]) % 2 * 2 - 1)
    blocks[3][9, 0] = np.nan                               # NaN row
    bb = BlockBatch(sizes)
    A0 = np.concatenate([b.ravel() for b in blocks])
    A = torch.from_numpy(A0.copy()).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    LUref, pivref, inforef, off, poff = ref_block_lu(sizes, blocks)
    assert inforef[0] == 8 and pivref[poff[2]] == 0 and pivref[poff[1] + 5] != 5
    assert (info.cpu().numpy() == inforef).all()
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    LU = A.cpu().numpy()
    fin = np.isfinite(LUref)
    assert (np.isfinite(LU) == fin).all()
    assert np.abs(LU[fin] - LUref[fin]).max() < 1e-12 * np.abs(LUref[fin]).max()
    # inverses: every block; the singular one keeps its factors and INFO
    A = torch.from_numpy(A0.copy()).to(bb.dev)
    info.zero_()
    bb.factor(A, None, info, inverse=True)
    torch.cuda.synchronize()
    inv = A.cpu().numpy()
    assert (info.cpu().numpy() == inforef).all()
    s0 = slice(off[0], off[0] + sizes[0] ** 2)
    assert np.abs(inv[s0] - LUref[s0]).max() < 1e-12 * np.abs(LUref[s0]).max()
    for b, n in enumerate(sizes):
        if b in (0, 3):
            continue
        Ainv = inv[off[b]:off[b] + n * n].reshape(n, n)
        assert np.abs(blocks[b] @ Ainv - np.eye(n)).max() < 1e-12 * n * max(1.0, np.abs(Ainv).max()), (b, n)
