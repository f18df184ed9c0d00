"""mife_box_build against the generic path on the same boxes, array by array and
bit for bit: coordinates and connectivity against the generated mesh, the
equations CSR against mife_csr_pattern, the Dirichlet reduction against
DirichletReduction (free DOFs ascending), and the SELL-64 storage, read back as a
reduced CSR (entries with a source, in order), against the reduced pattern and
its source positions; then a served request's solve through the box structures
against the oracle."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _build(ne, order, ncomp, lo, hi, extent, u_lo, u_hi):
    from iron_amd import _lib
    dev = "cuda"
    d = _lib.BoxDesc(tuple(ne), order, ncomp, lo, hi, 0, extent, u_lo, u_hi)
    c = _lib.BoxCounts()
    _lib.call("mife_box_sizes", C.byref(d), C.byref(c))
    t = dict(coords=torch.empty((c.n_nodes, 3), dtype=torch.float64, device=dev),
             elem_nodes=torch.empty((c.n_elem, c.nodes_per_elem), dtype=torch.int32, device=dev),
             row_ptr=torch.empty(c.n_dofs + 1, dtype=torch.int32, device=dev),
             col=torch.empty(c.nnz, dtype=torch.int32, device=dev),
             free_idx=torch.empty(c.n_free, dtype=torch.int32, device=dev),
             new_idx=torch.empty(c.n_dofs, dtype=torch.int32, device=dev),
             u0=torch.empty(c.n_dofs, dtype=torch.float64, device=dev),
             slice_ptr=torch.empty(c.n_slices + 1, dtype=torch.int32, device=dev),
             sell_src=torch.empty(c.stored, dtype=torch.int32, device=dev))
    if c.col16:
        t["sell_col16"] = torch.empty(c.stored, dtype=torch.int16, device=dev)
        t["col_base"] = torch.empty(c.n_slices, dtype=torch.int32, device=dev)
    else:
        t["sell_col"] = torch.empty(c.stored, dtype=torch.int32, device=dev)
    names = ("coords", "elem_nodes", "row_ptr", "col", "free_idx", "new_idx", "u0", "slice_ptr", "sell_col",
             "sell_col16", "col_base", "sell_src")
    arr = _lib.BoxArrays(*[_lib.ptr(t.get(n)) for n in names])
    _lib.call("mife_box_build", C.byref(d), C.byref(arr), _lib.stream_ptr())
    torch.cuda.synchronize()
    return c, {k: v.cpu().numpy() for k, v in t.items()}


@pytest.mark.parametrize("ne,order,ncomp,lo,hi", [((5, 3, 4), 1, 1, 1, 1), ((3, 2, 3), 2, 1, 1, 1),
                                                  ((6, 2, 3), 1, 3, 1, 0), ((1, 1, 1), 1, 3, 1, 0),
                                                  ((2, 1, 3), 2, 1, 0, 1), ((9, 8, 7), 1, 1, 1, 1),
                                                  ((7, 5, 6), 1, 3, 1, 0)])
def test_box_structures_equal_the_generic_path(ne, order, ncomp, lo, hi):
    from iron_amd import fem
    from iron_amd.basis import basis
    from iron_amd.mesh import DofMap, RegularMesh
    print(f"case box {ne} order {order} comps {ncomp} fix {lo}{hi}", flush=True)
    extent = (3.0, 1.0, 2.0)
    c, t = _build(ne, order, ncomp, lo, hi, extent, -0.5, 1.25)
    b = basis(3, order, order + 1)
    m = RegularMesh(ne, [b], extent)
    dm = DofMap(m, [0] * ncomp)
    assert np.array_equal(t["coords"], m.coords)
    assert np.array_equal(t["elem_nodes"], m.elem_nodes[0])
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, "cuda")
    rp, col = csr.to_host()
    assert np.array_equal(t["row_ptr"], rp) and np.array_equal(t["col"], col)
    N = m.num_nodes(0)
    ixn = m.node_grid_pos[:, 0]
    fixed_node = ((ixn == 0) & bool(lo)) | ((ixn == order * ne[0]) & bool(hi))
    fixed = np.tile(fixed_node, ncomp)
    red = fem.DirichletReduction(csr, fixed)
    assert np.array_equal(t["free_idx"], red.free_idx.cpu().numpy())
    assert np.array_equal(t["new_idx"], red.new_idx.cpu().numpy())
    u0 = np.where(np.tile((ixn == 0) & bool(lo), ncomp), -0.5, np.where(fixed, 1.25, 0.0))
    assert np.array_equal(t["u0"], u0)
    # SELL -> reduced CSR: slice s, lane l, entry j at slice_ptr[s] + 64 j + l
    rrp = red.row_ptr.cpu().numpy()
    rcol = red.col[: red.nnz].cpu().numpy()
    rsrc = red.src_pos[: red.nnz].cpu().numpy()
    W, ns = c.width, c.n_slices
    assert c.red_nnz == red.nnz
    assert np.array_equal(t["slice_ptr"], np.arange(ns + 1) * 64 * W)
    src = t["sell_src"].reshape(ns, W, 64)
    if c.col16:
        colv = (t["sell_col16"].astype(np.int64) & 0xFFFF).reshape(ns, W, 64) + t["col_base"][:, None, None]
    else:
        colv = t["sell_col"].reshape(ns, W, 64)
    for k in range(red.n_free):
        s, l = divmod(k, 64)
        sk, ck = src[s, :, l], colv[s, :, l]
        real = sk >= 0
        n_real = int(real.sum())
        assert np.all(real[:n_real]) and not np.any(real[n_real:])       # entries first, then padding
        assert np.array_equal(ck[:n_real], rcol[rrp[k]:rrp[k + 1]])
        assert np.array_equal(sk[:n_real], rsrc[rrp[k]:rrp[k + 1]])
        assert np.all(ck[n_real:] == k)                                   # padding: the row itself
    for k in range(red.n_free, ns * 64):                                  # rows past the end: padding
        s, l = divmod(k, 64)
        assert np.all(src[s, :, l] == -1)


def test_served_request_on_box_structures_matches_the_oracle():
    """One request of each kind through solve_one_fp64 (box structures, structured
    assembly, FP64 CG) against the oracle's direct solution."""
    from iron_amd import serving
    from test_gpu_serving import _oracle_solution
    for req in (serving.Request(0, "diffusion", (6, 5, 4), 21), serving.Request(1, "elasticity", (5, 3, 4), 22)):
        print(f"case served {req.kind} {req.n_elem}", flush=True)
        out = serving.solve_one_fp64(req, {})
        ref = _oracle_solution(req)
        assert out.relres <= 1e-8
        assert np.max(np.abs(out.u - ref)) <= 1e-6 * max(1.0, np.max(np.abs(ref)))
