"""The C restatement (oracle/fe_oracle.c, the CPU baseline's code) against the
NumPy oracle, which the reference's golden files and closed forms pin: assembly
(serial and OpenMP), the Dirichlet reduction and Jacobi CG."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import fe_oracle as fo

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def lib():
    so = ROOT / "oracle" / "liboracle_fe.so"
    if not so.exists():
        from oracle import build
        build.build()
    L = C.CDLL(str(so))
    L.oracle_reduce.restype = C.c_int64
    L.oracle_cg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                            C.c_int, C.c_int, C.POINTER(C.c_double)]
    return L


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _system(ne, order, seed=3):
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.5, 0.75))
    od = fo.DofMap(om, [0])
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.standard_normal(om.num_elements))
    return ob, om, od, kappa


@pytest.mark.parametrize("ne,order", [((3, 2, 2), 2), ((4, 3, 2), 1)])
def test_c_assembly_serial_and_threads_match_numpy(lib, ne, order):
    ob, om, od, kappa = _system(ne, order)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    ref = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    sref = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    dN = np.ascontiguousarray(ob.dxi_table(), dtype=np.float64)        # [ng][nn][3]
    N = np.ascontiguousarray(ob.value_table(), dtype=np.float64)       # [ng][nn]
    w = np.ascontiguousarray(ob.gauss_weights, dtype=np.float64)
    E, ng, nn = om.num_elements, dN.shape[0], dN.shape[1]
    coords = np.ascontiguousarray(om.coords, dtype=np.float64)
    en = np.ascontiguousarray(om.elem_nodes[0] - 1, dtype=np.int32)
    ed = np.ascontiguousarray(od.element_dofs(), dtype=np.int32)
    rp32, col32 = rp.astype(np.int32), col.astype(np.int32)
    f = np.ones(E)
    out = {}
    for mode in ("serial", "threads"):
        vals = np.zeros(len(col))
        src = np.zeros(od.num_dofs)
        args = (E, nn, ng, P(dN), P(N), P(w), P(coords), P(en), P(ed), P(kappa), P(f), P(rp32), P(col32), P(vals),
                P(src), 0, E)
        rc = lib.oracle_assemble_scalar3(*args) if mode == "serial" else lib.oracle_assemble_scalar3_par(*args, 4)
        assert rc == 0
        out[mode] = (vals, src)
        scale = np.abs(ref).max()
        assert np.abs(vals - ref).max() <= 1e-13 * scale, mode
        assert np.abs(src - sref).max() <= 1e-13 * np.abs(sref).max(), mode
    assert np.abs(out["serial"][0] - out["threads"][0]).max() <= 1e-14 * np.abs(ref).max()


def test_c_reduction_and_cg_match_numpy(lib):
    ob, om, od, kappa = _system((3, 3, 2), 2, seed=8)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    X = om.coords
    fixed = (X[:, 0] == 0.0) | (X[:, 0] == 1.0)
    ufix = np.where(X[:, 0] == 1.0, 2.0, 0.0)
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, src)
    new_idx = np.full(od.num_dofs, -1, dtype=np.int32)
    new_idx[free] = np.arange(len(free), dtype=np.int32)
    u = np.where(fixed, ufix, 0.0)
    n_f = len(free)
    srp_c = np.zeros(n_f + 1, dtype=np.int32)
    scol_c = np.zeros(len(scol), dtype=np.int32)
    sv_c = np.zeros(len(scol))
    b_c = np.zeros(n_f)
    nnz = lib.oracle_reduce(od.num_dofs, P(rp.astype(np.int32)), P(col.astype(np.int32)), P(vals), P(new_idx), P(u),
                            P(src), P(srp_c), P(scol_c), P(sv_c), P(b_c))
    assert nnz == len(scol)
    assert np.array_equal(srp_c, srp) and np.array_equal(scol_c, scol)
    assert np.array_equal(sv_c, sv) and np.abs(b_c - b).max() <= 1e-15 * np.abs(b).max()
    x = np.zeros(n_f)
    rel = C.c_double(0.0)
    for threads in (1, 4):
        x[:] = 0.0
        it = lib.oracle_cg(n_f, P(srp_c), P(scol_c), P(sv_c), P(b_c), P(x), 1e-12, 10000, threads, C.byref(rel))
        assert 0 < it < 10000 and rel.value <= 1e-12
        xd = np.linalg.solve(fo._csr_scipy(srp, scol, sv).toarray(), b)
        assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()


def _c_arrays(ob):
    dN = np.ascontiguousarray(ob.dxi_table(), dtype=np.float64)
    w = np.ascontiguousarray(ob.gauss_weights, dtype=np.float64)
    return dN, w


@pytest.mark.parametrize("threads", [1, 3])
def test_c_elasticity_assembly_matches_numpy(lib, threads):
    """Orthotropic per-element materials on a distorted trilinear mesh (iron's DXI_DX
    index order as written), serial and OpenMP."""
    ob = fo.Basis(3, 1, 2)
    ne, extent = (3, 2, 2), (3.0, 1.0, 1.0)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    rng = np.random.default_rng(6)
    X = om.coords.copy()
    inner = ~np.any((np.abs(X) < 1e-12) | (np.abs(X - np.array(extent)) < 1e-12), axis=1)
    X[inner] += 0.1 * rng.uniform(-1, 1, (int(inner.sum()), 3))
    E = om.num_elements
    mat = np.ascontiguousarray(np.stack([rng.uniform(5, 15, E), rng.uniform(3, 8, E), rng.uniform(1, 4, E),
                                         rng.uniform(0.05, 0.15, E), rng.uniform(0.2, 0.3, E),
                                         rng.uniform(0.3, 0.35, E)], axis=1))
    rp, col = fo.csr_pattern(od)
    ref = fo.assemble_csr(rp, col, od.element_dofs(),
                          fo.linear_elasticity_element_matrices(X[om.elem_nodes[0] - 1], ob,
                                                                fo.linear_elasticity_tensor(*mat.T)), od.num_dofs)
    dN, w = _c_arrays(ob)
    en = np.ascontiguousarray(om.elem_nodes[0] - 1, dtype=np.int32)
    ed = np.ascontiguousarray(od.element_dofs(), dtype=np.int32)
    rp32, col32 = rp.astype(np.int32), col.astype(np.int32)
    vals = np.zeros(len(col))
    Xc = np.ascontiguousarray(X)
    args = (E, ob.number_of_nodes, len(w), P(dN), P(w), P(Xc), P(en), P(ed), P(mat), P(rp32), P(col32), P(vals),
            0, E)
    rc = lib.oracle_assemble_elastic3(*args) if threads == 1 else lib.oracle_assemble_elastic3_par(*args, threads)
    assert rc == 0
    assert np.abs(vals - ref).max() <= 1e-13 * np.abs(ref).max()


def test_c_backsubstitute_matches_numpy(lib):
    ob, om, od, kappa = _system((3, 2, 2), 1)
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    u = np.random.default_rng(1).standard_normal(od.num_dofs)
    R = np.zeros(od.num_dofs)
    lib.oracle_backsubstitute(od.num_dofs, P(rp.astype(np.int32)), P(col.astype(np.int32)), P(vals), P(u), P(src),
                              P(R), 1)
    ref = fo._csr_scipy(rp, col, vals) @ u - src
    assert np.abs(R - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("order,elem_p", [(2, False), (1, True)])
def test_c_mooney_rivlin_element_loop_matches_numpy(lib, order, elem_p):
    """The C element loop (residual + exact Jacobian, deviatoric branch) against the
    NumPy oracle, which finite differences, closed forms and the reference's
    goldens pin (tests/test_mr_oracle.py)."""
    from oracle import mr_oracle as mo
    bu = fo.Basis(3, order, order + 1)
    bases = [bu] if elem_p else [bu, fo.Basis(3, 1, order + 1)]
    om = fo.regular_mesh((2, 2, 1), bases, (2.0, 1.5, 1.0))
    E = om.num_elements
    rng = np.random.default_rng(9)
    prob = mo.MRProblem(om, elem_p, rng.uniform(1, 3, E), rng.uniform(0.5, 6, E))
    x = prob.initial_state(-1.0)
    x[:3 * prob.nu] += 0.04 * rng.standard_normal(3 * prob.nu)
    x[3 * prob.nu:] += rng.standard_normal(prob.npd)
    R_o, _ = prob.residual(x)
    K_o = prob.jacobian(x)
    K_o.sum_duplicates()
    K_o.sort_indices()
    dN, w = _c_arrays(bu)
    psi = np.ones((len(w), 1)) if elem_p else np.ascontiguousarray(prob.bp.value_table(), dtype=np.float64)
    npd = psi.shape[1]
    en = np.ascontiguousarray(prob.en_u, dtype=np.int32)
    ed = np.ascontiguousarray(prob.edofs, dtype=np.int32)
    mat = np.ascontiguousarray(np.stack([prob.c1, prob.c2], axis=1))
    coords = np.ascontiguousarray(prob.geo)
    vals = np.zeros(len(K_o.indices))
    R = np.zeros(prob.ndofs)
    lib.oracle_mr_assemble.restype = C.c_int
    rc = lib.oracle_mr_assemble(E, bu.number_of_nodes, npd, len(w), prob.nu, P(dN), P(np.ascontiguousarray(psi)),
                                P(w), P(coords), P(np.ascontiguousarray(x)), P(en), P(ed), P(mat),
                                P(K_o.indptr.astype(np.int32)), P(K_o.indices.astype(np.int32)), P(vals), P(R), 0, E)
    assert rc == 0
    assert np.abs(R - R_o).max() <= 1e-12 * np.abs(R_o).max()
    assert np.abs(vals - K_o.data).max() <= 1e-12 * np.abs(K_o.data).max()


def test_c_gmres_matches_dense_solve(lib):
    """Restarted GMRES with a diagonal preconditioner (one override) on a
    non-symmetric system, against a dense solve."""
    rng = np.random.default_rng(2)
    n = 120
    A = np.diag(4.0 + rng.uniform(0, 1, n)) + 0.3 * rng.standard_normal((n, n)) * (rng.uniform(0, 1, (n, n)) < 0.05)
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    S.sort_indices()
    b = rng.standard_normal(n)
    pc = np.zeros(n)
    pc[5] = 0.2
    x = np.zeros(n)
    rr = C.c_double(0.0)
    lib.oracle_gmres.restype = C.c_int
    lib.oracle_gmres.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    it = lib.oracle_gmres(n, P(S.indptr.astype(np.int32)), P(S.indices.astype(np.int32)), P(S.data), P(pc), P(b),
                          P(x), 1e-12, 10, 1000, 1, C.byref(rr))
    assert 0 < it < 1000 and rr.value <= 1e-12
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10 * np.abs(x).max()
