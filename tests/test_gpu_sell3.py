"""Node-block SELL3 storage of a three-component solver matrix (3x3 blocks, one
16-bit node column per block) against the scalar storages: products, the Jacobi
diagonal through CG, FP64 CG and mixed FP32/FP64 CG, on trilinear elasticity
boxes clamped at x = 0 with a lognormal Young's modulus.  Fields whose free DOFs
are not whole nodes keep the scalar layout."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _elastic(ne, extent, seed=7, fixed_fn=None):
    from iron_amd import basis, engine, fem, mesh
    b = basis.basis(3, 1)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0, 0, 0])
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, colours=mesh.structured_colours(m))
    X = m.coords
    E, N = m.num_elements, m.num_nodes(0)
    g = torch.Generator().manual_seed(seed)
    Ee = 10.0 * torch.exp(0.3 * torch.randn(E, dtype=torch.float64, generator=g)).numpy()
    mat = np.stack([Ee, Ee, Ee, np.full(E, 0.3), np.full(E, 0.3), np.full(E, 0.3)], axis=1)
    v = csr.zeros()
    fem.ElasticAssembler(b, X, m.elem_nodes[0], csr).assemble(v, torch.as_tensor(mat, device="cuda"))
    if fixed_fn is None:
        clamp = X[:, 0] == 0.0
        fixed = np.concatenate([clamp, clamp, clamp])
    else:
        fixed = fixed_fn(X)
    red = fem.DirichletReduction(csr, fixed)
    tip = np.nonzero(X[:, 0] == extent[0])[0]
    f = torch.as_tensor(engine.neumann_integrate(m, dm, [b] * 3, {int(2 * N + i): -1.0 for i in tip}),
                        device="cuda")
    u = torch.zeros(csr.n, dtype=torch.float64, device="cuda")
    bvec = red.rhs(v, u, rhs=f)
    return fem, red, v, bvec, N


def _product(S, vals, x_orig, prec):
    """A x in the original row order through storage S."""
    from iron_amd import fem
    if S.fmt == "csr":
        y = torch.empty_like(x_orig)
        return S.spmv(vals, x_orig, y, prec)
    perm = S.perm.long()
    xs = x_orig[perm].contiguous()
    ys = torch.empty_like(xs)
    S.spmv(vals, xs, ys, prec)
    y = torch.empty_like(ys)
    y[perm] = ys
    return y


@pytest.mark.parametrize("schedule", [True, False])
def test_sell3_products_match_scalar_storages(schedule):
    from iron_amd import fem
    print(f"case SELL3 products 12x6x5 schedule={schedule}", flush=True)
    fem, red, v, _, N = _elastic((12, 6, 5), (2.0, 1.0, 1.0))
    S3 = fem.SolverMatrix(red, interleave=(3, N), block=True, schedule=schedule)
    Ss = fem.SolverMatrix(red, interleave=(3, N), schedule=schedule)
    Sc = fem.SolverMatrix(red, "csr")
    assert S3.block == 3 and S3.fmt == "sell3" and S3.col16 is not None
    assert Ss.block == 1
    # every solver entry is stored once; the rest is zero padding
    v3 = S3.gather(v)
    assert abs(float(v3.sum()) - float(Sc.gather(v).sum())) <= 1e-9 * float(v3.abs().sum())
    assert int((v3 != 0).sum()) <= red.nnz
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.rand(red.n_free, dtype=torch.float64, device="cuda", generator=g) - 0.5
    y3 = _product(S3, v3, x, fem.F64)
    ys = _product(Ss, Ss.gather(v), x, fem.F64)
    yc = _product(Sc, Sc.gather(v), x, fem.F64)
    scale = float(yc.abs().max())
    assert float((y3 - yc).abs().max()) <= 1e-12 * scale
    assert float((ys - yc).abs().max()) <= 1e-12 * scale
    # FP32 values and vectors
    x32 = x.float()
    y3f = _product(S3, S3.gather(v, fem.F32), x32, fem.F32)
    assert float((y3f.double() - yc).abs().max()) <= 2e-5 * scale


def test_sell3_cg_and_mixed_cg_match_scalar():
    from iron_amd import fem
    print("case SELL3 CG 24x8x8", flush=True)
    fem, red, v, bvec, N = _elastic((24, 8, 8), (3.0, 1.0, 1.0))
    S3 = fem.SolverMatrix(red, interleave=(3, N), block=True)
    Ss = fem.SolverMatrix(red, interleave=(3, N))
    assert S3.block == 3
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    sols = {}
    for name, S in (("sell3", S3), ("sell", Ss)):
        x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        info = fem.cg(S, S.gather(v), bvec, x, 1e-10, 100000, ws, prec=fem.F64)
        assert info.relres <= 1e-10
        xm = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        s64, s32 = S.gather2(v)
        im = fem.refine(S, s64, s32, bvec, xm, 1e-9, 100000, ws)
        assert im.relres <= 1e-9
        print(f"{name}: CG {info.iterations} its, mixed {im.iterations} inner its / {im.outer} replacements",
              flush=True)
        sols[name] = (x, xm, info.iterations)
    x3, xm3, it3 = sols["sell3"]
    xs, xms, its = sols["sell"]
    scale = float(xs.abs().max())
    assert float((x3 - xs).abs().max()) <= 1e-7 * scale
    assert float((xm3 - xs).abs().max()) <= 1e-6 * scale
    assert abs(it3 - its) <= max(3, its // 50)
    # the same storage solves the same system bit for bit again
    x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    fem.cg(S3, S3.gather(v), bvec, x, 1e-10, 100000, ws, prec=fem.F64)
    assert torch.equal(x, x3)


def test_sell3_gmres_and_minres_match_scalar():
    from iron_amd import fem
    print("case SELL3 GMRES / MINRES 12x4x4", flush=True)
    fem, red, v, bvec, N = _elastic((12, 4, 4), (3.0, 1.0, 1.0), seed=3)
    S3 = fem.SolverMatrix(red, interleave=(3, N), block=True)
    Ss = fem.SolverMatrix(red, interleave=(3, N))
    assert S3.block == 3
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    out = {}
    for name, S in (("sell3", S3), ("sell", Ss)):
        vals = S.gather(v)
        xg = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        ig = fem.gmres(S, vals, bvec, xg, 1e-10, 20000, ws, restart=30)
        xm = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        im = fem.minres(S, S.gather(v, fem.F32), bvec, xm, 1e-9, 20000, ws)
        print(f"{name}: GMRES {ig.iterations} its relres {ig.relres:.1e}, MINRES (FP32 values) {im.iterations} its "
              f"relres {im.relres:.1e}", flush=True)
        assert ig.relres <= 1e-10
        out[name] = (xg, xm, ig.iterations)
    scale = float(out["sell"][0].abs().max())
    assert float((out["sell3"][0] - out["sell"][0]).abs().max()) <= 1e-7 * scale
    # MINRES on the FP32 operator: both storages hold the same rounded values
    assert float((out["sell3"][1] - out["sell"][1]).abs().max()) <= 1e-5 * scale
    # restarted GMRES(30) follows the rounding of the products closely: within 5%
    assert abs(out["sell3"][2] - out["sell"][2]) <= max(3, out["sell"][2] // 20)


def test_sell3_declines_partial_nodes():
    from iron_amd import fem
    print("case SELL3 declined for rollers", flush=True)

    def rollers(X):
        # u_x = 0 on x = 0 only: those nodes keep two free components
        fx = X[:, 0] == 0.0
        fy = (X[:, 1] == 0.0) & (X[:, 0] == 0.0)
        fz = (X[:, 2] == 0.0) & (X[:, 0] == 0.0)
        return np.concatenate([fx, fy, fz])

    fem, red, v, _, N = _elastic((6, 4, 4), (1.0, 1.0, 1.0), fixed_fn=rollers)
    S = fem.SolverMatrix(red, interleave=(3, N), block=True)
    assert S.block == 1 and S.fmt == "sell"


def test_sell3_graph_replay_and_noalloc_bitwise():
    """On SELL3 the CG result is the same bit for bit whether its iteration chunks
    replay cached HIP graphs or launch kernel by kernel, and through the
    entry point that allocates nothing."""
    import ctypes as C
    from iron_amd import _lib
    print("case SELL3 graphs / no-alloc 24x8x8", flush=True)
    fem, red, v, bvec, N = _elastic((24, 8, 8), (3.0, 1.0, 1.0), seed=9)
    S = fem.SolverMatrix(red, interleave=(3, N), block=True)
    assert S.block == 3 and red.n_free <= 131072   # the graph path's size
    vals = S.gather(v)
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    xg = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    ig = fem.cg(S, vals, bvec, xg, 1e-10, 100000, ws, prec=fem.F64)
    _lib.graphs_enable(False)
    try:
        xe = torch.zeros_like(xg)
        ie = fem.cg(S, vals, bvec, xe, 1e-10, 100000, ws, prec=fem.F64)
    finally:
        _lib.graphs_enable(True)
    assert ig.iterations == ie.iterations and torch.equal(xg, xe)
    xn = torch.zeros_like(xg)
    it, rr = C.c_int(0), C.c_double(0.0)
    rc = _lib.lib().mife_cg_noalloc(C.byref(S.desc), _lib.ptr(vals), _lib.ptr(bvec), _lib.ptr(xn), fem.F64, 1e-10,
                                    100000, C.byref(it), C.byref(rr), _lib.stream_ptr())
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    assert it.value == ie.iterations and torch.equal(xn, xe)
