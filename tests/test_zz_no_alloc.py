"""mife_cg_noalloc: the CG entry point that allocates nothing (its buffers live in
a device array the library holds per GPU) against mife_cg with a caller's
workspace, through the C ABI: bit-equal solutions and equal iteration counts on
two calls in a row, the second on a smaller system (the array keeps the first
call's contents), in FP64 and in FP32.  Systems beyond the array's size are
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from oracle import fe_oracle as fo

# two solver states (524,416 bytes, 524,544 with alignment) and seven vectors of n
# values, each 256-byte aligned, in 64 MiB
FP64_MAX_ROWS = 1_179_616
FP32_MAX_ROWS = 2_359_232


def _desc(n):
    from iron_amd import _lib
    return _lib.MatrixDesc(n, _lib.MIFE_FMT_CSR, 1, 1, None, 0, 10, None, None, 0)


def test_size_limit_refused_before_device_work():
    """The documented row limits, and a null pointer, are checked on the host."""
    from iron_amd import _lib
    L = _lib.lib()
    it, rr = C.c_int(0), C.c_double(0.0)
    for n, prec in ((FP64_MAX_ROWS + 1, _lib.MIFE_F64), (FP32_MAX_ROWS + 1, _lib.MIFE_F32)):
        d = _desc(n)
        assert L.mife_cg_noalloc(C.byref(d), 1, 1, 1, prec, 1e-8, 10, C.byref(it), C.byref(rr), None) == 2
        assert "scratch" in _lib.last_error()
    assert L.mife_cg_noalloc(None, 1, 1, 1, 0, 1e-8, 10, None, None, None) == -1
    assert L.mife_cg_noalloc(C.byref(_desc(10)), 1, 1, 1, 3, 1e-8, 10, None, None, None) == 2


def _poisson(ne, seed):
    """Lognormal-conductivity trilinear Poisson, u = 0 on the boundary (oracle FP64)."""
    rng = np.random.default_rng(seed)
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    kappa = np.exp(rng.standard_normal(om.num_elements))
    xe = om.coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    X = om.coords
    fixed = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1.0) < 1e-12), axis=1)
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, np.zeros(od.num_dofs), src)
    return srp.astype(np.int32), scol.astype(np.int32), sv, b


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_matches_mife_cg_twice_second_smaller(prec):
    import torch
    from iron_amd import _lib
    L = _lib.lib()
    dt = torch.float64 if prec == "f64" else torch.float32
    p = _lib.MIFE_F64 if prec == "f64" else _lib.MIFE_F32
    for ne, seed in (((14, 12, 10), 1), ((7, 6, 5), 2)):
        print(f"case noalloc prec={prec} ne={ne}", flush=True)
        srp, scol, sv, b = _poisson(ne, seed)
        n = len(b)
        rp_d = torch.as_tensor(srp, device="cuda")
        col_d = torch.as_tensor(scol, device="cuda")
        v_d = torch.as_tensor(sv, device="cuda").to(dt)
        b_d = torch.as_tensor(b, device="cuda").to(dt)
        desc = _lib.MatrixDesc(n, _lib.MIFE_FMT_CSR, rp_d.data_ptr(), col_d.data_ptr(), None, 0, len(scol), None,
                               None, 0)
        assert rp_d.shape[0] == n + 1 and col_d.shape[0] == int(srp[-1]) and v_d.shape[0] == int(srp[-1])
        rtol = 1e-10 if prec == "f64" else 1e-5
        wb = int(L.mife_solver_work_bytes(n, 1))
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        x_ref = torch.zeros(n, dtype=dt, device="cuda")
        it_r, rr_r = C.c_int(0), C.c_double(0.0)
        rc = L.mife_cg(C.byref(desc), v_d.data_ptr(), b_d.data_ptr(), x_ref.data_ptr(), p, rtol, 5000,
                       work.data_ptr(), wb, C.byref(it_r), C.byref(rr_r), stream)
        assert rc == 0, _lib.last_error()
        x_na = torch.zeros(n, dtype=dt, device="cuda")
        it_n, rr_n = C.c_int(0), C.c_double(0.0)
        rc = L.mife_cg_noalloc(C.byref(desc), v_d.data_ptr(), b_d.data_ptr(), x_na.data_ptr(), p, rtol, 5000,
                               C.byref(it_n), C.byref(rr_n), stream)
        assert rc == 0, _lib.last_error()
        torch.cuda.synchronize()
        assert it_n.value == it_r.value and it_n.value > 5
        assert rr_n.value == rr_r.value
        assert torch.equal(x_na, x_ref)
        A = fo._csr_scipy(srp, scol, sv)
        xh = x_na.double().cpu().numpy()
        tol = 1.01 * rtol if prec == "f64" else 1e-4
        assert np.linalg.norm(b - A @ xh) / np.linalg.norm(b) <= tol
