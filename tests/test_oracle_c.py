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
