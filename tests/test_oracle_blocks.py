"""Oracle block LU against LAPACK getrf (scipy.linalg.lu_factor): same pivot sequence, same factors."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle.bosss_ref import ref_block_lu, ref_block_solve


@pytest.mark.parametrize("n", [1, 3, 10, 20, 50])
def test_getf2_matches_lapack(n):
    rng = np.random.default_rng(n)
    blocks = [rng.uniform(-1, 1, (n, n)) for _ in range(5)]
    LU, piv, info, off, poff = ref_block_lu([n] * 5, blocks)
    assert (info == 0).all()
    for b in range(5):
        lu, p = sla.lu_factor(blocks[b])
        assert (piv[poff[b]:poff[b] + n] == p).all()
        assert np.abs(LU[off[b]:off[b] + n * n].reshape(n, n) - lu).max() < 1e-12 * max(1, np.abs(lu).max())
    x = rng.uniform(-1, 1, 5 * n)
    xs = ref_block_solve([n] * 5, LU, piv, off, poff, x)
    for b in range(5):
        assert np.allclose(blocks[b] @ xs[b * n:(b + 1) * n], x[b * n:(b + 1) * n], atol=1e-11)


def test_zero_pivot_info():
    A = np.array([[1.0, 2.0], [2.0, 4.0]])
    _, _, info, _, _ = ref_block_lu([2], [A])
    assert info[0] == 2
