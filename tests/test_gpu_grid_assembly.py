"""The structured (grid) assembly against the coloured scatter and the oracle.

mife_assemble_grid stores the element matrices in one launch and writes every
CSR row whole, summing the element entries of an entry in the colour order of the
coloured scatter: the values must equal the coloured scatter's bit for bit (FP64
and FP32, with the Poisson source vector), on boxes and on moved interior nodes,
and the oracle's to 1e-13 (FP64)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402


def _mesh(ne, order, ncomp, extent, distort):
    from iron_amd import basis, mesh
    b = basis.basis(3, {1: 1, 2: 2}[order], order + 1)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0] * ncomp)
    X = m.coords.copy()
    if distort:
        rng = np.random.default_rng(3)
        h = np.asarray(extent) / np.asarray(ne) / order
        inner = ~np.any((np.abs(X) < 1e-12) | (np.abs(X - np.asarray(extent)) < 1e-12), axis=1)
        X[inner] += 0.2 * h * rng.uniform(-1, 1, (int(inner.sum()), 3))
    return b, m, dm, X


@pytest.mark.parametrize("ne,order,distort", [((5, 4, 3), 1, False), ((3, 4, 2), 2, False), ((4, 3, 3), 1, True),
                                              ((3, 2, 3), 2, True), ((1, 1, 1), 2, False)])
def test_scalar_grid_equals_coloured_scatter(ne, order, distort):
    from iron_amd import fem, mesh
    print(f"case grid scalar ne={ne} order={order} distort={distort}", flush=True)
    extent = (1.0, 1.5, 0.75)
    b, m, dm, X = _mesh(ne, order, 1, extent, distort)
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, colours=mesh.structured_colours(m))
    col_asm = fem.ScalarAssembler(b, X, m.elem_nodes[0], csr)
    grid_asm = fem.ScalarAssembler(b, X, m.elem_nodes[0], csr, grid=ne)
    assert grid_asm.grid is not None and col_asm.grid is None
    rng = np.random.default_rng(8)
    kt = torch.as_tensor(np.exp(rng.standard_normal(m.num_elements)), device="cuda")
    ft = torch.as_tensor(rng.uniform(0.5, 2.0, m.num_elements), device="cuda")
    for prec in (fem.F64, fem.F32):
        v_c, v_g = csr.zeros(prec), csr.zeros(prec)
        s_c = torch.zeros(dm.num_dofs, dtype=torch.float64, device="cuda")
        s_g = torch.full_like(s_c, 7.0)          # overwritten, not added to
        v_g.fill_(3.0)
        col_asm.assemble(v_c, kappa=kt, source=ft, src_vec=s_c, prec=prec)
        grid_asm.assemble(v_g, kappa=kt, source=ft, src_vec=s_g, prec=prec)
        torch.cuda.synchronize()
        assert torch.equal(v_g[: csr.nnz], v_c[: csr.nnz])
        assert torch.equal(s_g, s_c)
    # against the oracle
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    xe = X[om.elem_nodes[0] - 1]
    ref = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kt.cpu().numpy()),
                          od.num_dofs)
    v = csr.zeros(fem.F64)
    grid_asm.assemble(v, kappa=kt, prec=fem.F64)
    assert np.abs(v.cpu().numpy()[: len(ref)] - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("distort", [False, True])
def test_elastic_grid_equals_coloured_scatter(distort):
    from iron_amd import fem, mesh
    print(f"case grid elastic distort={distort}", flush=True)
    ne, extent = (4, 3, 2), (3.0, 1.0, 1.0)
    b, m, dm, X = _mesh(ne, 1, 3, extent, distort)
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, colours=mesh.structured_colours(m))
    col_asm = fem.ElasticAssembler(b, X, m.elem_nodes[0], csr)
    grid_asm = fem.ElasticAssembler(b, X, m.elem_nodes[0], csr, grid=ne)
    rng = np.random.default_rng(2)
    E = m.num_elements
    mat = np.stack([rng.uniform(5, 15, E), rng.uniform(3, 8, E), rng.uniform(1, 4, E), rng.uniform(0.05, 0.15, E),
                    rng.uniform(0.2, 0.3, E), rng.uniform(0.3, 0.35, E)], axis=1)
    mt = torch.as_tensor(np.ascontiguousarray(mat), device="cuda")
    for prec in (fem.F64, fem.F32):
        v_c, v_g = csr.zeros(prec), csr.zeros(prec)
        col_asm.assemble(v_c, mt, prec=prec)
        grid_asm.assemble(v_g, mt, prec=prec)
        torch.cuda.synchronize()
        assert torch.equal(v_g[: csr.nnz], v_c[: csr.nnz])
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    rp, col = fo.csr_pattern(od)
    ref = fo.assemble_csr(rp, col, od.element_dofs(),
                          fo.linear_elasticity_element_matrices(X[om.elem_nodes[0] - 1], ob,
                                                                fo.linear_elasticity_tensor(*mat.T)), od.num_dofs)
    v = csr.zeros(fem.F64)
    grid_asm.assemble(v, mt, prec=fem.F64)
    assert np.abs(v.cpu().numpy()[: len(ref)] - ref).max() <= 1e-13 * np.abs(ref).max()
