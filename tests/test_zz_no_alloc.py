"""bosss_hip_block_lu_noalloc_dev, the allocation-free batched block LU, called through the C ABI
and checked against bosss_hip_block_lu_dev on the same blocks: two calls in a row (the second with
fewer and smaller blocks, so stale scratch from the first would show), large blocks (split form,
its per-block status words in the library's fixed device array), mid-size blocks (the existing
entry point factors them one workgroup per block, the no-alloc one in the split form: the same
panels, substitutions and tile products), small blocks with inverse (register kernels), zero
pivots, and the
refusal of a batch beyond the scratch size before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def batch(sizes, seed, zero_cols=()):
    rng = np.random.default_rng(seed)
    blocks = [rng.uniform(-1, 1, (n, n)) for n in sizes]
    for b, c in zero_cols:
        blocks[b][:, c] = 0.0                      # a zero pivot: INFO = c + 1, factorisation continues
    n = np.array(sizes, np.int32)
    off = np.zeros(len(n), np.int64); off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    poff = np.zeros(len(n), np.int64); poff[1:] = np.cumsum(n.astype(np.int64))[:-1]
    return n, off, poff, np.concatenate([b.ravel() for b in blocks])


def factor(entry, n, off, poff, A, inverse=False):
    import torch
    from bosss_amd._lib import lib
    from bosss_amd.operator import Context
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    n_t, off_t, poff_t, A_t = t(n), t(off), t(poff), t(A)
    ipiv = torch.zeros(int(n.sum()), dtype=torch.int32, device=dev)
    info = torch.full((len(n),), -7, dtype=torch.int32, device=dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    L = lib()
    rc = getattr(L, entry)(Context.get(0).h, len(n), int(n.max()), p(n_t), p(off_t), p(A_t), None if inverse else p(ipiv),
                           None if inverse else p(poff_t), p(info), 1 if inverse else 0,
                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L.bosss_hip_block_last_error().decode()
    torch.cuda.synchronize()
    return A_t.cpu().numpy(), ipiv.cpu().numpy(), info.cpu().numpy()


CALLS = [
    ("large, split form in both", [600] * 12 + [531] * 5, [(3, 0), (7, 301)]),
    ("mid-size, smaller second call", [200] * 3 + [129], [(1, 57)]),
]


@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_noalloc_lu_matches_block_lu_dev_on_consecutive_calls(order):
    for i in order:
        name, sizes, zc = CALLS[i]
        print("case noalloc block LU:", name, len(sizes), "blocks", flush=True)
        n, off, poff, A = batch(sizes, 11 + i, zc)
        ref = factor("bosss_hip_block_lu_dev", n, off, poff, A)
        got = factor("bosss_hip_block_lu_noalloc_dev", n, off, poff, A)
        assert (got[2] == ref[2]).all(), (got[2], ref[2])
        assert (got[2][[b for b, _ in zc]] == [c + 1 for _, c in zc]).all()
        assert (got[1] == ref[1]).all()
        assert np.abs(got[0] - ref[0]).max() <= 1e-12 * np.abs(ref[0]).max()


def test_noalloc_small_blocks_with_inverse():
    print("case noalloc block LU: small blocks with inverse", flush=True)
    sizes = [10] * 40 + [20] * 6
    n, off, poff, A = batch(sizes, 5)
    ref = factor("bosss_hip_block_lu_dev", n, off, poff, A, inverse=True)
    got = factor("bosss_hip_block_lu_noalloc_dev", n, off, poff, A, inverse=True)
    assert (got[2] == ref[2]).all() and (got[2] == 0).all()
    assert np.array_equal(got[0], ref[0])


def test_noalloc_refuses_batches_beyond_its_scratch():
    import torch
    from bosss_amd._lib import lib
    from bosss_amd.operator import Context
    print("case noalloc block LU: 1025 blocks of 130 refused", flush=True)
    dev = torch.device("cuda:0")
    nb = 1025
    n_t = torch.full((nb,), 130, dtype=torch.int32, device=dev)
    off_t = torch.zeros(nb, dtype=torch.int64, device=dev)
    A_t = torch.zeros(130 * 130, dtype=torch.float64, device=dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    L = lib()
    rc = L.bosss_hip_block_lu_noalloc_dev(Context.get(0).h, nb, 130, p(n_t), p(off_t), p(A_t), None, None, None, 0,
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1
    assert "scratch" in L.bosss_hip_block_last_error().decode()
    torch.cuda.synchronize()
