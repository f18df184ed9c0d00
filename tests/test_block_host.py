"""Joint solver pattern of several equations sets (fem.BlockReduction) on the host:
rows and columns set by set in iron's order (solver_mapping_routines.F90:791-830),
block diagonal, equal to scipy's block_diag of the parts' patterns."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp

torch = pytest.importorskip("torch")


def _part(A):
    A = sp.csr_matrix(A)
    return SimpleNamespace(csr=SimpleNamespace(device=torch.device("cpu")), n_free=A.shape[0], nnz=A.nnz,
                           row_ptr=torch.as_tensor(A.indptr.astype(np.int32)),
                           col=torch.as_tensor(A.indices.astype(np.int32)), A=A)


def test_block_pattern_is_the_block_diagonal_of_the_parts():
    from iron_amd import fem
    rng = np.random.default_rng(0)
    mats = []
    for n in (5, 1, 7):
        M = (rng.random((n, n)) < 0.4).astype(float)
        np.fill_diagonal(M, 1.0)
        mats.append(M)
    parts = [_part(M) for M in mats]
    blk = fem.BlockReduction(parts)
    ref = sp.block_diag([p.A for p in parts], format="csr")
    assert blk.n_free == ref.shape[0] and blk.nnz == ref.nnz
    assert np.array_equal(blk.row_ptr.numpy(), ref.indptr)
    assert np.array_equal(blk.col.numpy()[: blk.nnz], ref.indices)
    assert blk.row_off == [0, 5, 6, 13] and blk.nnz_off[-1] == ref.nnz
    assert np.array_equal(blk.src_pos.numpy(), np.arange(ref.nnz))
