"""Serving path (grouped CG in FP64 and in FP32 with FP64 reliable updates,
HBM-budget admission, continuous batching, HIP-graph replay): every request against
the FP64 oracle and the single-solve path, a request alone against the same request
in a mix (bit for bit), and admission under a tight budget."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402


def _oracle_solution(req):
    """Dense FP64 solve of the request's reduced system with the oracle's assembly."""
    ob = fo.Basis(3, 1, 2)
    extent = (1.0, 1.0, 1.0) if req.kind == "diffusion" else (3.0, 1.0, 1.0)
    om = fo.regular_mesh(req.n_elem, [ob], extent)
    ncomp = 1 if req.kind == "diffusion" else 3
    od = fo.DofMap(om, [0] * ncomp)
    ed = od.element_dofs()
    xe = om.coords[om.elem_nodes[0] - 1]
    E = om.num_elements
    from iron_amd import serving
    z = serving.lognormal_field(E, req.seed, "cuda").cpu().numpy()
    X = om.coords
    N = len(X)
    if req.kind == "diffusion":
        Ke = fo.laplace_element_matrices(xe, ob, z)
        fixed = (X[:, 0] == 0.0) | (X[:, 0] == 1.0)
        ufix = np.where(X[:, 0] == 1.0, 1.0, 0.0)
        rhs = np.zeros(N)
    else:
        C = fo.linear_elasticity_tensor(10 * z, 10 * z, 10 * z, np.full(E, 0.3), np.full(E, 0.3), np.full(E, 0.3))
        Ke = fo.linear_elasticity_element_matrices(xe, ob, C)
        clamp = X[:, 0] == 0.0
        fixed = np.concatenate([clamp] * 3)
        ufix = np.zeros(3 * N)
        ny, nz = req.n_elem[1], req.n_elem[2]
        rhs = np.zeros(3 * N)
        tip = np.nonzero(X[:, 0] == 3.0)[0]
        iy = np.rint(X[tip, 1] * ny).astype(int)
        iz = np.rint(X[tip, 2] * nz).astype(int)
        cy = np.where((iy == 0) | (iy == ny), 0.5, 1.0)
        cz = np.where((iz == 0) | (iz == nz), 0.5, 1.0)
        rhs[2 * N + tip] = -1.0 * cy * cz / (ny * nz)
    rp, col = fo.csr_pattern(od)
    vals = fo.assemble_csr(rp, col, ed, Ke, od.num_dofs)
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, rhs)
    A = fo._csr_scipy(srp, scol, sv).toarray()
    u = np.where(fixed, ufix, 0.0)
    u[free] = np.linalg.solve(A, b)
    return u


def _mix():
    from iron_amd import serving
    return [serving.Request(0, "diffusion", (4, 3, 5), 11), serving.Request(1, "elasticity", (6, 2, 2), 12),
            serving.Request(2, "diffusion", (7, 7, 6), 13), serving.Request(3, "elasticity", (4, 3, 3), 14),
            serving.Request(4, "diffusion", (3, 3, 3), 15)]


@pytest.mark.parametrize("use_graph,precision", [(False, "fp64"), (True, "fp64"), (False, "mixed"), (True, "auto"),
                                                (True, "split")])
def test_grouped_requests_match_oracle_and_single_solve(use_graph, precision):
    from iron_amd import serving
    print(f"case serving mix graph={use_graph} precision={precision}", flush=True)
    srv = serving.Server(hbm_budget=2e8, max_segments=16, chunk=8, use_graph=use_graph, precision=precision)
    reqs = srv.run(_mix())
    single = {}
    for r in reqs:
        assert r.relres <= 1e-8, (r.rid, r.relres)
        ref = _oracle_solution(r)
        assert np.abs(r.u - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max()), r.rid
        one = serving.solve_one_fp64(serving.Request(r.rid, r.kind, r.n_elem, r.seed), single)
        assert np.abs(one.u - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max())
    if use_graph:
        assert srv.stats["graph_replays"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alone_equals_in_mix_bitwise(precision):
    """The target request (elasticity) iterates in FP64 or FP32; the mix around it
    holds requests of both precisions (policy "split") or all FP32."""
    from iron_amd import serving
    print(f"case serving alone vs mix precision={precision}", flush=True)
    target = serving.Request(9, "elasticity", (5, 3, 4), 99)
    other = "split" if precision == "fp64" else "mixed"
    alone = serving.Server(hbm_budget=2e8, max_segments=16, chunk=8, precision=precision).run(
        [serving.Request(9, "elasticity", (5, 3, 4), 99)])[0]
    mix = _mix() + [target]
    for k, r in enumerate(mix):
        r.arrival = 0.0 if k % 2 == 0 else 1e-3  # joins at different chunk boundaries
    out = serving.Server(hbm_budget=2e8, max_segments=16, chunk=8, precision=other).run(mix)
    inmix = [r for r in out if r.rid == 9][0]
    assert inmix.iterations == alone.iterations
    assert np.array_equal(inmix.u, alone.u)


def test_budget_admission_queues_and_completes():
    from iron_amd import serving
    print("case serving tight budget", flush=True)
    reqs = serving.make_stream(12, seed=5, lo=3, hi=8)
    srv = serving.Server(hbm_budget=3e6, max_segments=4, chunk=4, use_graph=False)
    out = srv.run(reqs)
    assert srv.stats["retired"] == 12
    for r in out:
        assert r.u is not None and r.relres <= 1e-8
    assert srv.stats["table_builds"] > 3


def test_mixed_and_fp64_servers_agree_and_mixed_stops_on_the_fp64_residual():
    """FP32 iterations with FP64 reliable updates reach the same 1e-8 FP64 residual
    as FP64 iterations (the reported residual is the true FP64 one), on a mix whose
    larger requests need several replacements."""
    from iron_amd import serving
    print("case serving mixed vs FP64", flush=True)
    reqs = serving.make_stream(6, seed=31, lo=6, hi=14)
    out = {}
    for prec in ("fp64", "mixed"):
        srv = serving.Server(hbm_budget=4e8, max_segments=16, chunk=16, precision=prec)
        out[prec] = {r.rid: r for r in srv.run([serving.Request(r.rid, r.kind, r.n_elem, r.seed) for r in reqs])}
    for rid, r64 in out["fp64"].items():
        r32 = out["mixed"][rid]
        assert r32.relres <= 1e-8 and r64.relres <= 1e-8
        ref = _oracle_solution(r64)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(r32.u - ref).max() <= 1e-6 * scale, rid
        assert np.abs(r32.u - r64.u).max() <= 1e-6 * scale, rid
        print(f"request {rid} {r64.kind} {r64.n_elem}: iterations FP32+FP64 updates {r32.iterations}, "
              f"FP64 {r64.iterations}", flush=True)
        if r64.kind == "diffusion":   # FP32 iterations track the FP64 ones on diffusion
            assert r32.iterations <= 1.25 * r64.iterations + 8, (rid, r32.iterations, r64.iterations)


@pytest.mark.parametrize("kind,ne", [("diffusion", (64, 64, 64)), ("elasticity", (64, 48, 40)),
                                     ("diffusion", (8, 64, 13))])
def test_served_solution_at_full_size_has_the_residual_it_reports(kind, ne):
    """Requests at the stream's largest sizes (beyond the dense oracle): the served
    solution u, put through an independent FP64 path -- the generic structures
    (mesh, DOF map, CSR pattern, position-map assembly, Dirichlet reduction) and
    the CSR SpMV -- has the relative residual ||b_f - K_ff u_f|| / ||b_f|| the
    server reports (<= 1e-8), its fixed DOFs hold the Dirichlet values, and the
    elasticity reactions balance the applied load."""
    from iron_amd import fem, serving
    from iron_amd.basis import basis
    from iron_amd.mesh import DofMap, RegularMesh
    print(f"case served full size {kind} {ne}", flush=True)
    req = serving.Request(0, kind, ne, 77)
    out = serving.Server(hbm_budget=8e9, max_segments=4, chunk=16).run([req])[0]
    assert out.relres <= 1e-8
    b = basis(3, 1, 2)
    extent = (1.0, 1.0, 1.0) if kind == "diffusion" else (3.0, 1.0, 1.0)
    m = RegularMesh(ne, [b], extent)
    nc = 1 if kind == "diffusion" else 3
    dm = DofMap(m, [0] * nc)
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, "cuda")
    st = serving._Structure(kind, ne, "cuda")
    coef = st.coefficients(req.seed, "cuda")
    A = fem.ScalarAssembler if kind == "diffusion" else fem.ElasticAssembler
    vals = csr.zeros()
    if kind == "diffusion":
        A(b, m.coords, m.elem_nodes[0], csr).assemble(vals, kappa=coef)
    else:
        A(b, m.coords, m.elem_nodes[0], csr).assemble(vals, coef)
    X = m.coords
    N = len(X)
    if kind == "diffusion":
        fixed = (X[:, 0] == 0.0) | (X[:, 0] == 1.0)
        ufix = np.where(X[:, 0] == 1.0, 1.0, 0.0)
    else:
        fixed = np.tile(X[:, 0] == 0.0, 3)
        ufix = np.zeros(3 * N)
    assert np.array_equal(out.u[fixed], ufix[fixed])
    rp, col = csr.to_host()
    u = torch.as_tensor(out.u, device="cuda")
    Ku = torch.zeros(csr.n, dtype=torch.float64, device="cuda")
    rows = torch.repeat_interleave(torch.arange(csr.n, device="cuda"), torch.as_tensor(np.diff(rp), device="cuda"))
    Ku.index_add_(0, rows, vals[: csr.nnz] * u[torch.as_tensor(col, device="cuda")])
    f = st.rhs if st.rhs is not None else torch.zeros(csr.n, dtype=torch.float64, device="cuda")
    free = torch.as_tensor(~fixed, device="cuda")
    # b_f = f_f - K_fd u_d, K_ff u_f = K u - K_fd u_d on the free rows
    ud = torch.as_tensor(np.where(fixed, ufix, 0.0), device="cuda")
    Kud = torch.zeros(csr.n, dtype=torch.float64, device="cuda")
    Kud.index_add_(0, rows, vals[: csr.nnz] * ud[torch.as_tensor(col, device="cuda")])
    bf = (f - Kud)[free]
    rf = (f - Ku)[free]
    rel = float(torch.linalg.vector_norm(rf) / torch.linalg.vector_norm(bf))
    print(f"served {kind} {ne}: reported {out.relres:.3e}, independent {rel:.3e}, {out.iterations} its", flush=True)
    assert rel <= 1.05e-8
    assert abs(rel - out.relres) <= 0.05 * out.relres + 1e-12
    if kind == "elasticity":
        # reactions at the clamp balance the applied z load (-1)
        R = (Ku - f)[torch.as_tensor(fixed, device="cuda")]
        Rz = float(R[2 * int(fixed.sum()) // 3:].sum())
        assert abs(Rz - 1.0) <= 1e-6
