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


@pytest.mark.parametrize("sizes", [[150, 140], [600, 520]])   # one-workgroup form; two launches per column
def test_large_block_zero_pivot_and_inverse_limit(sizes):
    import torch
    from bosss_amd.blocks import BlockBatch
    say("case large block singular", sizes)
    blocks = make_blocks(sizes, 9)
    blocks[0][:, 37] = 0.0                  # zero column -> zero pivot at j = 37 (INFO = 38)
    bb = BlockBatch(sizes)
    A = torch.from_numpy(np.concatenate([b.ravel() for b in blocks])).to(bb.dev)
    ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=bb.dev)
    info = torch.zeros(bb.nblk, dtype=torch.int32, device=bb.dev)
    bb.factor(A, ipiv, info)
    torch.cuda.synchronize()
    _, pivref, inforef, _, _ = ref_block_lu(sizes, blocks)
    assert list(info.cpu().numpy()) == [38, 0] == list(inforef)
    assert (ipiv.cpu().numpy() == pivref + 1).all()
    with pytest.raises(RuntimeError):
        bb.factor(A, None, info, inverse=True)   # inverses are limited to 100 DOF
