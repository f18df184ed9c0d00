"""mife_box_sizes: the closed-form sizes of a generated box's structures (host only,
no device work) against the oracle's pattern and Dirichlet reduction of the same
regular mesh, and the argument checks."""
import ctypes as C

import numpy as np
import pytest

from oracle import fe_oracle as fo


def _sizes(ne, order, ncomp, lo, hi):
    from iron_amd import _lib
    d = _lib.BoxDesc(tuple(ne), order, ncomp, lo, hi, 0, (1.0, 1.0, 1.0), 0.0, 1.0)
    c = _lib.BoxCounts()
    rc = _lib.lib().mife_box_sizes(C.byref(d), C.byref(c))
    return rc, c


@pytest.mark.parametrize("ne,order,ncomp,lo,hi", [((3, 2, 2), 1, 1, 1, 1), ((2, 3, 2), 2, 1, 1, 0),
                                                  ((3, 2, 4), 1, 3, 1, 0), ((1, 1, 1), 1, 3, 1, 0),
                                                  ((4, 1, 2), 2, 1, 0, 1), ((2, 2, 1), 1, 1, 0, 0)])
def test_box_sizes_match_the_oracle_pattern(ne, order, ncomp, lo, hi):
    rc, c = _sizes(ne, order, ncomp, lo, hi)
    assert rc == 0
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0] * ncomp)
    rp, col = fo.csr_pattern(od)
    N = len(om.coords)
    ix = np.rint(om.coords[:, 0] * order * ne[0]).astype(int)
    fixed_node = ((ix == 0) & bool(lo)) | ((ix == order * ne[0]) & bool(hi))
    fixed = np.tile(fixed_node, ncomp)
    free = np.nonzero(~fixed)[0]
    # reduced pattern: free rows, free columns
    red = [np.count_nonzero(~fixed[col[rp[r]:rp[r + 1]]]) for r in free]
    assert c.n_nodes == N and c.n_elem == np.prod(ne) and c.n_dofs == ncomp * N
    assert c.nnz == rp[-1]
    assert c.n_free == len(free)
    assert c.red_nnz == sum(red)
    assert c.width == max(red)
    assert c.n_slices == -(-len(free) // 64) and c.stored == c.n_slices * 64 * c.width
    assert c.nodes_per_elem == (order + 1) ** 3


@pytest.mark.parametrize("args", [((2, 2, 2), 3, 1, 1, 1), ((2, 2, 2), 1, 2, 1, 1), ((0, 2, 2), 1, 1, 1, 1),
                                  ((1, 2, 2), 1, 1, 1, 1), ((2, 2, 2), 1, 1, 2, 0)])
def test_box_sizes_reject_bad_boxes(args):
    rc, _ = _sizes(*args)
    assert rc != 0
