"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerances: CSR pattern, DOF maps and scatter positions bit-exact; FP64
assembly within 1e-13 of the largest entry (summation order differs); FP32
assembly within 2e-6 (FP32 contraction on FP64 geometry); solutions within the
stated relative residual of the FP64 system.
"""
import json
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"


def _fem():
    from iron_amd import fem, mesh, basis  # noqa: F401
    return fem, mesh, basis


def _build(ne, order, ncomp=1, extent=None, gauss=None):
    fem, mesh, basis = _fem()
    d = len(ne)
    extent = extent or (1.0,) * d
    gx = gauss or order + 1
    pb = basis.basis(d, {1: 1, 2: 2}[order], gx)
    pm = mesh.RegularMesh(ne, [pb], extent)
    pd = mesh.DofMap(pm, [0] * ncomp)
    ob = fo.Basis(nxi=d, order=order, ngauss_xi=gx)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0] * ncomp)
    return pb, pm, pd, ob, om, od


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("ne,order,ncomp", [((3, 2, 2), 1, 1), ((2, 3, 2), 2, 1), ((3, 2, 2), 1, 3), ((4, 3), 1, 1),
                                           ((3, 2), 2, 1)])
def test_pattern_and_dofs_bit_exact(ne, order, ncomp):
    fem, _, _ = _fem()
    pb, pm, pd, ob, om, od = _build(ne, order, ncomp)
    assert np.array_equal(pd.element_dofs(), od.element_dofs())
    csr = fem.CSRStructure(pd.element_dofs(), pd.num_dofs)
    rp, col = csr.to_host()
    orp, ocol = fo.csr_pattern(od)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    # scatter map: position of (row i, col j) of every element entry
    ed = od.element_dofs()
    pos = csr.pos.cpu().numpy().reshape(len(ed), ed.shape[1], ed.shape[1])
    for e in [0, len(ed) // 2, len(ed) - 1]:
        for i in range(ed.shape[1]):
            for j in range(ed.shape[1]):
                assert pos[e, i, j] == fo.storage_location_find(orp, ocol, ed[e, i], ed[e, j])
    # colouring: no DOF shared inside a colour
    order_ = csr.order.cpu().numpy()
    for c in range(csr.n_colours):
        els = order_[csr.colour_ptr[c]:csr.colour_ptr[c + 1]]
        dofs = pd.element_dofs()[els].ravel()
        assert len(np.unique(dofs)) == len(dofs)


@pytest.mark.parametrize("ne,order", [((4, 3, 3), 1), ((3, 2, 3), 2), ((5, 4), 1), ((3, 3), 2)])
def test_scalar_assembly_fp64_fp32(ne, order):
    fem, _, _ = _fem()
    rng = np.random.default_rng(7)
    pb, pm, pd, ob, om, od = _build(ne, order, extent=tuple(1.0 + 0.5 * k for k in range(len(ne))))
    kappa = np.exp(rng.standard_normal(pm.num_elements))
    csr = fem.CSRStructure(pd.element_dofs(), pd.num_dofs)
    asm = fem.ScalarAssembler(pb, pm.coords, pm.elem_nodes[0], csr)
    kt = torch.as_tensor(kappa, device="cuda")
    ft = torch.full((pm.num_elements,), 1.0, dtype=torch.float64, device="cuda")
    v64 = csr.zeros(fem.F64)
    s64 = torch.zeros(pd.num_dofs, dtype=torch.float64, device="cuda")
    asm.assemble(v64, kappa=kt, source=ft, src_vec=s64, prec=fem.F64)
    v32 = csr.zeros(fem.F32)
    asm.assemble(v32, kappa=kt, prec=fem.F32)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    ref = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    sref = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    assert _rel(v64.cpu().numpy()[: len(ref)], ref) < 1e-13
    assert _rel(s64.cpu().numpy(), sref) < 1e-13
    assert _rel(v32.cpu().numpy()[: len(ref)].astype(np.float64), ref) < 2e-6
    # determinism of the coloured scatter and agreement with the atomic path
    v64b = csr.zeros(fem.F64)
    asm.assemble(v64b, kappa=kt, prec=fem.F64)
    assert torch.equal(v64, v64b)
    va = csr.zeros(fem.F64)
    asm.assemble(va, kappa=kt, prec=fem.F64, atomic=True)
    assert _rel(va.cpu().numpy()[: len(ref)], ref) < 1e-13


def test_elastic_assembly_fp64_fp32():
    fem, _, _ = _fem()
    rng = np.random.default_rng(3)
    pb, pm, pd, ob, om, od = _build((4, 2, 3), 1, ncomp=3, extent=(3.0, 1.0, 1.5))
    E = 10.0 * np.exp(0.3 * rng.standard_normal(pm.num_elements))
    mat = np.stack([E, E, E, 0.3 * np.ones_like(E), 0.3 * np.ones_like(E), 0.3 * np.ones_like(E)], axis=1)
    csr = fem.CSRStructure(pd.element_dofs(), pd.num_dofs)
    asm = fem.ElasticAssembler(pb, pm.coords, pm.elem_nodes[0], csr)
    mt = torch.as_tensor(np.ascontiguousarray(mat), device="cuda")
    v64 = csr.zeros(fem.F64)
    asm.assemble(v64, mt, prec=fem.F64)
    v32 = csr.zeros(fem.F32)
    asm.assemble(v32, mt, prec=fem.F32)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    C = fo.linear_elasticity_tensor(*mat.T)
    ref = fo.assemble_csr(rp, col, od.element_dofs(), fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs)
    assert _rel(v64.cpu().numpy()[: len(ref)], ref) < 1e-13
    assert _rel(v32.cpu().numpy()[: len(ref)].astype(np.float64), ref) < 2e-6


def _poisson_system(ne, order, seed=11):
    """Lognormal-kappa Poisson with f=1 and u=0 on every boundary face (oracle)."""
    rng = np.random.default_rng(seed)
    pb, pm, pd, ob, om, od = _build(ne, order)
    kappa = np.exp(rng.standard_normal(pm.num_elements))
    X = om.coords
    fixed = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1.0) < 1e-12), axis=1)
    return pb, pm, pd, ob, om, od, kappa, fixed


def _gpu_linear(pb, pm, pd, kappa, fixed, ufix, source=1.0):
    fem, _, _ = _fem()
    csr = fem.CSRStructure(pd.element_dofs(), pd.num_dofs)
    asm = fem.ScalarAssembler(pb, pm.coords, pm.elem_nodes[0], csr)
    kt = torch.as_tensor(kappa, device="cuda")
    ft = torch.full((pm.num_elements,), float(source), dtype=torch.float64, device="cuda")
    v64 = csr.zeros(fem.F64)
    src = torch.zeros(pd.num_dofs, dtype=torch.float64, device="cuda")
    asm.assemble(v64, kappa=kt, source=ft if source else None, src_vec=src if source else None, prec=fem.F64)
    v32 = csr.zeros(fem.F32)
    asm.assemble(v32, kappa=kt, prec=fem.F32)
    red = fem.DirichletReduction(csr, fixed)
    u = torch.as_tensor(np.where(fixed, ufix, 0.0), device="cuda")
    b = red.rhs(v64, u, src=src if source else None)
    return fem, csr, red, v64, v32, src, u, b


def test_reduction_rhs_and_cg_fp64_against_oracle():
    pb, pm, pd, ob, om, od, kappa, fixed = _poisson_system((4, 3, 3), 2)
    ufix = np.zeros(pd.num_dofs)
    fem, csr, red, v64, v32, src, u, b = _gpu_linear(pb, pm, pd, kappa, fixed, ufix)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    sref = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    free, srp, scol, sv, bref = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, sref)
    assert np.array_equal(red.row_ptr.cpu().numpy(), srp) and np.array_equal(red.col.cpu().numpy()[: len(scol)], scol)
    sv_gpu = fem.SolverMatrix(red, "csr").gather(v64, fem.F64)
    assert _rel(sv_gpu.cpu().numpy()[: len(sv)], sv) < 1e-13
    assert _rel(b.cpu().numpy(), bref) < 1e-13
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    xo, _, _ = fo.cg_jacobi(srp, scol, sv, bref, rtol=1e-13)
    A = fo._csr_scipy(srp, scol, sv)
    for fmt in ("csr", "sell"):
        S = fem.SolverMatrix(red, fmt)
        s64 = S.gather(v64, fem.F64)
        x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        info = fem.cg(S, s64, b, x, 1e-10, 10000, ws, prec=fem.F64)
        assert info.relres <= 1e-10, fmt
        assert _rel(x.cpu().numpy(), xo) < 1e-8, fmt
        # mixed precision: FP32 CG with FP64 residual replacement, to 1e-8 of the FP64 system
        s32 = S.gather(v32, fem.F32)
        xm = torch.zeros_like(x)
        info_m = fem.refine(S, s64, s32, b, xm, 1e-8, 10000, ws)
        assert np.linalg.norm(bref - A @ xm.cpu().numpy()) / np.linalg.norm(bref) <= 1e-8, fmt
        assert info_m.relres <= 1e-8 and info_m.outer >= 2, fmt
        # GMRES(30) + Jacobi
        xg = torch.zeros_like(x)
        info_g = fem.gmres(S, s64, b, xg, 1e-10, 20000, ws)
        assert info_g.relres <= 1e-10, fmt
        assert _rel(xg.cpu().numpy(), xo) < 1e-7, fmt


@pytest.mark.parametrize("compress", [False, True])
def test_sell_storage_matches_csr_product(compress):
    """SELL-64 values/columns (int32 or 16-bit slice-relative) reproduce the CSR solver
    matrix product in FP64, and CG on either storage converges to the same solution."""
    fem, _, _ = _fem()
    pb, pm, pd, ob, om, od, kappa, fixed = _poisson_system((5, 4, 3), 2, seed=5)
    fem, csr, red, v64, v32, src, u, b = _gpu_linear(pb, pm, pd, kappa, fixed, np.zeros(pd.num_dofs))
    Sc = fem.SolverMatrix(red, "csr")
    Ss = fem.SolverMatrix(red, "sell", sigma=128, compress=compress)
    assert (Ss.col16 is not None) == compress
    if compress:
        cb = Ss.col_base.cpu().numpy()
        c16 = Ss.col16.cpu().numpy().view(np.uint16).astype(np.int64)
        sp = Ss.slice_ptr.cpu().numpy()
        full = np.concatenate([cb[k] + c16[sp[k]:sp[k + 1]] for k in range(len(sp) - 1)])
        assert np.array_equal(full, Ss.col[: Ss.stored].cpu().numpy())
    vc = Sc.gather(v64, fem.F64)
    vs = Ss.gather(v64, fem.F64)
    # both precisions in one gather: the FP64 copy and its FP32 rounding, bit for bit
    for S_ in (Sc, Ss):
        d64, d32 = S_.gather2(v64)
        assert torch.equal(d64, S_.gather(v64, fem.F64))
        assert torch.equal(d32, S_.gather(v64, fem.F32))
        assert torch.equal(d32, d64.float())
    xv = torch.rand(red.n_free, dtype=torch.float64, device="cuda")
    yc = Sc.spmv(vc, xv, torch.empty_like(xv))
    perm = Ss.perm.long()
    ys = Ss.spmv(vs, xv[perm].contiguous(), torch.empty_like(xv))
    y_back = torch.empty_like(xv)
    y_back[perm] = ys
    assert _rel(y_back.cpu().numpy(), yc.cpu().numpy()) < 1e-14
    assert Ss.stored >= red.nnz and Ss.stored % 64 == 0
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    xc = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    xs = torch.zeros_like(xc)
    ic = fem.cg(Sc, vc, b, xc, 1e-12, 10000, ws)
    is_ = fem.cg(Ss, vs, b, xs, 1e-12, 10000, ws)
    assert ic.relres <= 1e-12 and is_.relres <= 1e-12
    assert _rel(xs.cpu().numpy(), xc.cpu().numpy()) < 1e-10


@pytest.mark.parametrize("dims", [(5, 4, 3), (12, 11, 10)])
def test_sell_schedule_covers_every_slice_and_keeps_products(dims):
    """The SELL schedule lists every slice once, by decreasing length inside each of
    its 8 contiguous ranges, each slot carrying the slice's first entry, entries per
    row and base column; products over the schedule (where several wavefronts share
    a slice on small matrices) agree with the plain slice order to rounding."""
    fem, _, _ = _fem()
    pb, pm, pd, ob, om, od, kappa, fixed = _poisson_system(dims, 2, seed=9)
    fem, csr, red, v64, v32, src, u, b = _gpu_linear(pb, pm, pd, kappa, fixed, np.zeros(pd.num_dofs))
    Sp = fem.SolverMatrix(red, "sell", schedule=False)
    Ss = fem.SolverMatrix(red, "sell")
    ns = Ss.desc.n_slices
    span = Ss.desc.order_span
    rng = -(-ns // 8)                      # slices per range, ceil(ns / 8)
    assert span == 4 * -(-rng // 4)
    slots = Ss.slice_order.cpu().numpy().reshape(8, span, 4)
    order = slots[:, :, 0]
    sp = Ss.slice_ptr.cpu().numpy()
    lens = np.diff(sp)
    cbase = Ss.col_base.cpu().numpy() if Ss.col_base is not None else np.zeros(ns, dtype=np.int32)
    seen = []
    for k in range(8):
        ids = order[k][order[k] >= 0]
        assert np.all(order[k][len(ids):] == -1)
        assert np.all(slots[k, len(ids):, 1:] == 0)
        assert np.array_equal(slots[k, :len(ids), 1], sp[ids])
        assert np.array_equal(slots[k, :len(ids), 2], lens[ids] // 64)
        assert np.array_equal(slots[k, :len(ids), 3], cbase[ids])
        assert np.array_equal(np.sort(ids), np.arange(min(k * rng, ns), min((k + 1) * rng, ns)))
        L = lens[ids]
        assert np.all(L[:-1] >= L[1:])
        seen.append(ids)
    assert np.array_equal(np.sort(np.concatenate(seen)), np.arange(ns))
    for prec, dt in ((fem.F64, torch.float64), (fem.F32, torch.float32)):
        vals = Ss.gather(v64, prec)
        xv = torch.rand(red.n_free, dtype=dt, device="cuda")
        yp = Sp.spmv(vals, xv, torch.empty_like(xv), prec)
        ys = Ss.spmv(vals, xv, torch.empty_like(xv), prec)
        tol = 1e-14 if prec == fem.F64 else 1e-6
        assert float((yp - ys).abs().max() / yp.abs().max()) <= tol
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    xs = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    info = fem.cg(Ss, Ss.gather(v64), b, xs, 1e-12, 10000, ws)
    assert info.relres <= 1e-12


def test_unit_cube_laplace_u_equals_x_and_reactions():
    fem, _, _ = _fem()
    pb, pm, pd, ob, om, od = _build((6, 5, 4), 1)
    X = pm.coords
    fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1.0) < 1e-12)
    ufix = np.where(np.abs(X[:, 0] - 1.0) < 1e-12, 1.0, 0.0)
    fem, csr, red, v64, v32, src, u, b = _gpu_linear(pb, pm, pd, np.ones(pm.num_elements), fixed, ufix, source=0.0)
    S = fem.SolverMatrix(red)
    sv = S.gather(v64, fem.F64)
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    info = fem.cg(S, sv, b, x, 1e-10, 10000, ws)
    red.field_update(x, u)
    un = u.cpu().numpy()
    assert np.max(np.abs(un - X[:, 0])) < 1e-8 and info.relres <= 1e-10
    R = fem.backsubstitute(csr, v64, u).cpu().numpy()
    assert abs(R[np.abs(X[:, 0]) < 1e-12].sum() + 1.0) < 1e-8
    assert abs(R[np.abs(X[:, 0] - 1) < 1e-12].sum() - 1.0) < 1e-8


def test_golden_analytic_laplace_through_gpu():
    """The reference's AnalyticLaplaceLinearLagrange.opanal reproduced by the HIP path."""
    fem, mesh, basis = _fem()
    data = json.loads((GOLDEN / "analytic_laplace_linear_lagrange.json").read_text())
    o = -data["setup"]["origin_pi"] / 2.0
    pb = basis.basis(2, 1, 2)
    pm = mesh.RegularMesh((2, 2), [pb], (2.0, 2.0), (o, o))
    pd = mesh.DofMap(pm, [0])
    X = pm.coords
    ua = np.cos(X[:, 0]) * np.cosh(X[:, 1])
    fixed = np.ones(9, bool)
    fixed[4] = False
    fem, csr, red, v64, v32, src, u, b = _gpu_linear(pb, pm, pd, np.ones(4), fixed, ua, source=0.0)
    S = fem.SolverMatrix(red)
    sv = S.gather(v64, fem.F64)
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    fem.cg(S, sv, b, x, 1e-14, 100, ws)
    red.field_update(x, u)
    un = u.cpu().numpy()
    for node, (num, _) in data["nodal"]["Phi"].items():
        assert float(f"{un[int(node) - 1]:.4e}") == pytest.approx(num, rel=1e-4, abs=1e-15)
    R = fem.backsubstitute(csr, v64, u, iron_compat=True).cpu().numpy()
    for node, (num, _) in data["nodal"]["del Phi/del n"].items():
        i = int(node) - 1
        if abs(num) < 1e-10:
            assert abs(R[i]) < 1e-12
        else:
            assert R[i] == pytest.approx(num, rel=1e-4)
    rp, col = fo.csr_pattern(fo.DofMap(fo.regular_mesh((2, 2), [fo.Basis(2, 1, 2)], (2.0, 2.0), (o, o)), [0]))
    vals = v64.cpu().numpy()[: len(col)]
    Rt = fem.backsubstitute(csr, v64, u, iron_compat=False).cpu().numpy()
    assert np.allclose(Rt, fo.backsubstitute(rp, col, vals, un), atol=1e-13)


@pytest.mark.parametrize("sigma", [1, 64, 100, 1024])
def test_sell_permutation_is_the_window_sort(sigma):
    """mife_sell_permutation against a sort of the composite key (window, decreasing
    length, position), with rows in their own order and at permuted positions."""
    from iron_amd import _lib
    print(f"case sell permutation window {sigma}", flush=True)
    rng = np.random.default_rng(sigma)
    n = 5000 + sigma // 3
    lengths = rng.choice([27, 45, 75, 125, 3, 0], size=n)
    row_ptr = torch.as_tensor(np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32), device="cuda")
    for permuted in (False, True):
        row_at = rng.permutation(n) if permuted else np.arange(n)
        pos = np.empty(n, dtype=np.int64)
        pos[row_at] = np.arange(n)
        L = lengths.max() + 1
        key = (pos // sigma) * (L * sigma) + (L - 1 - lengths) * sigma + pos % sigma
        want = np.argsort(key, kind="stable")
        perm = torch.empty(n, dtype=torch.int32, device="cuda")
        ra = torch.as_tensor(row_at.astype(np.int32), device="cuda") if permuted else None
        _lib.call("mife_sell_permutation", n, row_ptr.data_ptr(), None if ra is None else ra.data_ptr(), sigma,
                  perm.data_ptr(), _lib.stream_ptr())
        assert np.array_equal(perm.cpu().numpy(), want), permuted
