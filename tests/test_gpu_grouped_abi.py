"""The grouped solver through its C ABI, as a caller outside the server would use
it: four systems in one group, one per combination of storage (CSR; SELL-64 with
int32 columns; SELL-64 with 16-bit columns) and precision (FP64 CG; FP32 CG with
FP64 reliable updates), iterated in chunks until every segment reports done, each
solution against the FP64 oracle's dense solve."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402


def _system(ne, seed):
    """Lognormal-kappa Poisson box (f = 1, u = 0 / 1 on x = 0 / 1): GPU reduction and
    the oracle's dense solution of the same reduced system."""
    from iron_amd import basis, fem, mesh
    b = basis.basis(3, 1, 2)
    m = mesh.RegularMesh(ne, [b], (1.0, 1.0, 1.0))
    dm = mesh.DofMap(m, [0])
    X = m.coords
    fixed = (X[:, 0] == 0.0) | (X[:, 0] == 1.0)
    ufix = np.where(X[:, 0] == 1.0, 1.0, 0.0)
    kappa = np.exp(np.random.default_rng(seed).standard_normal(m.num_elements))
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, colours=mesh.structured_colours(m))
    asm = fem.ScalarAssembler(b, X, m.elem_nodes[0], csr)
    v = csr.zeros()
    src = torch.zeros(dm.num_dofs, dtype=torch.float64, device="cuda")
    ones = torch.ones(m.num_elements, dtype=torch.float64, device="cuda")
    asm.assemble(v, kappa=torch.as_tensor(kappa, device="cuda"), source=ones, src_vec=src)
    red = fem.DirichletReduction(csr, fixed)
    u = torch.as_tensor(np.where(fixed, ufix, 0.0), device="cuda")
    rhs = red.rhs(v, u, src=src)
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    vals = fo.assemble_csr(rp, col, od.element_dofs(),
                           fo.laplace_element_matrices(om.coords[om.elem_nodes[0] - 1], ob, kappa), od.num_dofs)
    sref = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(om.coords[om.elem_nodes[0] - 1], ob, 1.0),
                              od.num_dofs)
    free, srp, scol, sv, bref = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, sref)
    xref = np.linalg.solve(fo._csr_scipy(srp, scol, sv).toarray(), bref)
    return red, v, rhs, xref


def test_grouped_abi_all_storages_and_precisions():
    from iron_amd import _lib, fem
    print("case grouped ABI: CSR / SELL int32 / SELL 16-bit x FP64 / mixed", flush=True)
    L = _lib.lib()
    cases = [((5, 4, 3), "csr", False, fem.F64), ((6, 3, 4), "csr", False, fem.F32),
             ((4, 7, 3), "sell", False, fem.F64), ((7, 5, 5), "sell", True, fem.F32)]
    rp_pool, col_pool, c16_pool, cb_pool, val_pool, b_list = [], [], [], [], [], []
    offs = dict(row=0, rp=0, nz=0, cb=0)
    segs, ref = [], []
    for slot, (ne, fmt, compress, prec) in enumerate(cases):
        red, v, rhs, xref = _system(ne, slot + 3)
        S = fem.SolverMatrix(red, fmt, compress=compress, schedule=False)
        vals = S.gather(v)
        n = red.n_free
        if fmt == "csr":
            rp, col, nzs = red.row_ptr, red.col[:red.nnz], red.nnz
            perm = None
            ns, cb_off = 0, -1
        else:
            rp, nzs, ns = S.slice_ptr[: S.desc.n_slices + 1], S.stored, S.desc.n_slices
            perm = S.perm.long()
            if compress:
                assert S.col16 is not None
                col, cb_off = None, offs["cb"]
                c16_pool.append((offs["nz"], S.col16[:nzs]))
                cb_pool.append(S.col_base[:ns])
            else:
                col, cb_off = S.col[:nzs], -1
        if col is not None:
            col_pool.append((offs["nz"], col))
        rp_pool.append(rp)
        val_pool.append(vals[:nzs])
        b_list.append(rhs[perm] if perm is not None else rhs)
        segs.append(_lib.Segment(n, 100000, slot, _lib.MIFE_FMT_CSR if fmt == "csr" else _lib.MIFE_FMT_SELL,
                                 offs["row"], offs["rp"], offs["nz"], 1e-10, ns, prec, cb_off))
        ref.append((xref, perm))
        offs["row"] += n
        offs["rp"] += len(rp)
        offs["nz"] += nzs
        offs["cb"] += ns if cb_off >= 0 else 0   # base columns are kept for 16-bit segments only
    dev = "cuda"
    nz_total, rows = offs["nz"], offs["row"]
    colp = torch.zeros(nz_total, dtype=torch.int32, device=dev)
    c16p = torch.zeros(nz_total, dtype=torch.int16, device=dev)
    for o, c in col_pool:
        colp[o:o + len(c)] = c
    for o, c in c16_pool:
        c16p[o:o + len(c)] = c
    rpp = torch.cat(rp_pool)
    cbp = torch.cat(cb_pool) if cb_pool else torch.zeros(1, dtype=torch.int32, device=dev)
    valp = torch.cat(val_pool)
    val32 = valp.float()
    # every segment's ranges inside the pools (checked here, before any kernel reads them)
    for sgm in segs:
        nrp = sgm.n + 1 if sgm.format == _lib.MIFE_FMT_CSR else sgm.n_slices + 1
        assert sgm.rp_off + nrp <= len(rpp) and sgm.row_off + sgm.n <= rows
        nzs = int(rpp[sgm.rp_off + nrp - 1].item())
        assert sgm.nz_off + nzs <= nz_total == len(valp)
        if sgm.cb_off >= 0:
            assert sgm.cb_off + sgm.n_slices <= len(cbp)
    pools = _lib.Pools(rpp.data_ptr(), colp.data_ptr(), c16p.data_ptr(), cbp.data_ptr(), valp.data_ptr(),
                       val32.data_ptr())
    b = torch.cat(b_list)
    x = torch.zeros(rows, dtype=torch.float64, device=dev)
    vec5 = torch.zeros(5 * rows, dtype=torch.float64, device=dev)
    vec6 = torch.zeros(6 * rows, dtype=torch.float32, device=dev)
    nseg = len(segs)
    max_blocks = sum(L.mife_grouped_blocks(s.n) for s in segs) + 2 * nseg
    state = torch.zeros(int(L.mife_grouped_state_bytes(nseg)), dtype=torch.uint8, device=dev)
    wb = int(L.mife_grouped_work_bytes(nseg, max_blocks))
    table = torch.zeros(wb, dtype=torch.uint8, device=dev)
    stage = np.zeros(wb, dtype=np.uint8)
    arr = (_lib.Segment * nseg)(*segs)
    nb = C.c_int(0)
    st = _lib.stream_ptr()
    _lib.call("mife_grouped_setup", arr, nseg, nseg, max_blocks, C.byref(nb), stage.ctypes.data, table.data_ptr(), st)
    _lib.call("mife_grouped_cg_init", nseg, nb.value, nseg, max_blocks, C.byref(pools), b.data_ptr(), x.data_ptr(),
              vec5.data_ptr(), vec6.data_ptr(), rows, table.data_ptr(), state.data_ptr(), st)
    it = np.zeros(nseg, dtype=np.int32)
    done = np.zeros(nseg, dtype=np.int32)
    rel = np.zeros(nseg)
    for _ in range(2000):
        _lib.call("mife_grouped_cg_iterate", nseg, nb.value, nseg, max_blocks, C.byref(pools), b.data_ptr(),
                  x.data_ptr(), vec5.data_ptr(), vec6.data_ptr(), rows, 16, 4, 0.1, table.data_ptr(),
                  state.data_ptr(), st)
        _lib.call("mife_grouped_cg_status", nseg, state.data_ptr(), it.ctypes.data, done.ctypes.data,
                  rel.ctypes.data, st)
        if done.all():
            break
    assert done.all(), (it, rel)
    _lib.call("mife_grouped_cg_residual", nseg, nb.value, nseg, max_blocks, C.byref(pools), b.data_ptr(),
              x.data_ptr(), table.data_ptr(), state.data_ptr(), st)
    _lib.call("mife_grouped_cg_status", nseg, state.data_ptr(), it.ctypes.data, done.ctypes.data, rel.ctypes.data, st)
    xs = x.cpu().numpy()
    for k, (seg, (xref, perm)) in enumerate(zip(segs, ref)):
        assert rel[k] <= 1e-10, (k, rel[k])
        xk = xs[seg.row_off:seg.row_off + seg.n]
        if perm is not None:
            out = np.empty_like(xk)
            out[perm.cpu().numpy()] = xk
            xk = out
        assert np.abs(xk - xref).max() <= 1e-8 * max(1.0, np.abs(xref).max()), k
