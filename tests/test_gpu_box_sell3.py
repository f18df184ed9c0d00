"""mife_box_build with block = 1 (three components): the free DOFs numbered node
by node and the reduced matrix as SELL3, against the same box built with the
scalar layout (which tests/test_gpu_box.py checks against the generic path):
the same equations CSR, the node-by-node numbering, and the reduced product
y = A x through both storages on random equations-matrix values."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _build(ne, order, lo, hi, block):
    from iron_amd import _lib
    dev = "cuda"
    d = _lib.BoxDesc(tuple(ne), order, 3, lo, hi, block, (3.0, 1.0, 2.0), 0.0, 0.5)
    c = _lib.BoxCounts()
    _lib.call("mife_box_sizes", C.byref(d), C.byref(c))
    t = dict(coords=None,
             elem_nodes=torch.empty((c.n_elem, c.nodes_per_elem), dtype=torch.int32, device=dev),
             row_ptr=torch.empty(c.n_dofs + 1, dtype=torch.int32, device=dev),
             col=torch.empty(c.nnz, dtype=torch.int32, device=dev),
             free_idx=torch.empty(c.n_free, dtype=torch.int32, device=dev),
             new_idx=torch.empty(c.n_dofs, dtype=torch.int32, device=dev),
             u0=torch.empty(c.n_dofs, dtype=torch.float64, device=dev),
             slice_ptr=torch.empty(c.n_slices + 1, dtype=torch.int32, device=dev),
             sell_src=torch.empty(c.stored, dtype=torch.int32, device=dev))
    if c.col16:
        t["sell_col16"] = torch.empty(c.stored // 9 if c.block else c.stored, dtype=torch.int16, device=dev)
        t["col_base"] = torch.empty(c.n_slices, dtype=torch.int32, device=dev)
    else:
        t["sell_col"] = torch.empty(c.stored, dtype=torch.int32, device=dev)
    names = ("coords", "elem_nodes", "row_ptr", "col", "free_idx", "new_idx", "u0", "slice_ptr", "sell_col",
             "sell_col16", "col_base", "sell_src")
    arr = _lib.BoxArrays(*[_lib.ptr(t.get(n)) for n in names])
    _lib.call("mife_box_build", C.byref(d), C.byref(arr), _lib.stream_ptr())
    torch.cuda.synchronize()
    return c, t


def _product(c, t, fmt, vals_eq, x):
    from iron_amd import _lib
    vals = torch.empty(c.stored, dtype=torch.float64, device="cuda")
    _lib.call("mife_gather_values", c.stored, _lib.ptr(t["sell_src"]), _lib.ptr(vals_eq), 0, _lib.ptr(vals), 0,
              _lib.stream_ptr())
    perm = torch.arange(c.n_free, dtype=torch.int32, device="cuda")
    colp = _lib.ptr(t.get("sell_col16")) if c.col16 else _lib.ptr(t["sell_col"])
    desc = _lib.MatrixDesc(int(c.n_free), fmt, _lib.ptr(t["slice_ptr"]), colp, _lib.ptr(perm), int(c.n_slices),
                           int(c.stored), _lib.ptr(t.get("col_base")), None, 0)
    y = torch.empty_like(x)
    _lib.call("mife_spmv", C.byref(desc), _lib.ptr(vals), _lib.ptr(x), _lib.ptr(y), 0, _lib.stream_ptr())
    torch.cuda.synchronize()
    return y, vals


@pytest.mark.parametrize("ne,order,lo,hi", [((6, 2, 3), 1, 1, 0), ((7, 5, 6), 1, 1, 0), ((4, 3, 3), 1, 1, 1),
                                            ((1, 1, 1), 1, 1, 0), ((2, 2, 2), 2, 1, 0)])
def test_box_sell3_equals_scalar_box(ne, order, lo, hi):
    from iron_amd import _lib
    print(f"case box SELL3 {ne} order {order} fix {lo}{hi}", flush=True)
    c0, t0 = _build(ne, order, lo, hi, 0)
    c1, t1 = _build(ne, order, lo, hi, 1)
    assert c0.block == 0 and c1.block == 1 and c1.col16 == 1
    assert c1.n_free == c0.n_free and c1.red_nnz == c0.red_nnz
    assert c1.n_slices == (c1.n_free // 3 + 63) // 64 and c1.stored % 9 == 0
    for k in ("row_ptr", "col", "elem_nodes", "u0"):
        assert torch.equal(t0[k], t1[k]), k
    n_dofs, nf = int(c0.n_dofs), int(c0.n_free) // 3
    N = n_dofs // 3
    ni0, ni1 = t0["new_idx"].cpu().numpy(), t1["new_idx"].cpu().numpy()
    free = ni0 >= 0
    comp = np.arange(n_dofs) // N
    assert np.array_equal(ni1 >= 0, free)
    assert np.array_equal(ni1[free], 3 * (ni0[free] - comp[free] * nf) + comp[free])
    fi1 = t1["free_idx"].cpu().numpy()
    assert np.array_equal(ni1[fi1], np.arange(3 * nf))
    # every block value's source is an equations entry of its own row
    src = t1["sell_src"].cpu().numpy()
    assert int((src >= 0).sum()) == int(c1.red_nnz)
    assert len(np.unique(src[src >= 0])) == int(c1.red_nnz)
    g = torch.Generator(device="cuda").manual_seed(5)
    vals_eq = torch.rand(int(c0.nnz), dtype=torch.float64, device="cuda", generator=g) - 0.5
    x0 = torch.rand(int(c0.n_free), dtype=torch.float64, device="cuda", generator=g) - 0.5
    fd = torch.as_tensor(np.nonzero(free)[0], device="cuda")
    i0 = torch.as_tensor(ni0, device="cuda").long()[fd]
    i1 = torch.as_tensor(ni1, device="cuda").long()[fd]
    x1 = torch.empty_like(x0)
    x1[i1] = x0[i0]
    y0, _ = _product(c0, t0, _lib.MIFE_FMT_SELL, vals_eq, x0)
    y1, _ = _product(c1, t1, _lib.MIFE_FMT_SELL3, vals_eq, x1)
    scale = float(y0.abs().max())
    assert float((y1[i1] - y0[i0]).abs().max()) <= 1e-12 * scale


@pytest.mark.parametrize("policy", ["fp64", "mixed"])
def test_served_stream_sell3_equals_scalar_layout(policy):
    """A small served stream (diffusion and elasticity boxes) with the elasticity
    boxes in SELL3 node blocks against the same stream on scalar SELL: every
    request's solution, iteration count and true residual."""
    from iron_amd import serving
    print(f"case served stream SELL3 vs SELL, policy {policy}", flush=True)
    res = {}
    for block in (True, False):
        serving._Structure.block = block
        try:
            srv = serving.Server(hbm_budget=4e9, max_segments=64, chunk=16, use_graph=False, precision=policy)
            out = srv.run(serving.make_stream(12, seed=77, lo=6, hi=12))
        finally:
            serving._Structure.block = True
        torch.cuda.synchronize()
        res[block] = {(r.kind, r.n_elem, r.seed): r for r in out}
    assert any(k[0] == "elasticity" for k in res[True])
    for k, rb in res[True].items():
        rs = res[False][k]
        assert rb.relres <= 1e-8 and rs.relres <= 1e-8
        scale = max(float(np.abs(rs.u).max()), 1e-300)
        if k[0] == "diffusion":
            # untouched by the layout, and independent of the group: bit for bit
            assert rb.iterations == rs.iterations and np.array_equal(rb.u, rs.u), k
        elif policy == "fp64":
            # the same FP64 CG up to the order of each row's sums
            assert float(np.abs(rb.u - rs.u).max()) <= 1e-9 * scale, k
            assert abs(rb.iterations - rs.iterations) <= max(3, rs.iterations // 20), k
        else:
            # FP32 recurrences on the rounded operator take their own paths: both meet
            # the FP64 residual bound above and agree to the conditioning's reach
            assert float(np.abs(rb.u - rs.u).max()) <= 1e-4 * scale, k
