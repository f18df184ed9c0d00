"""HIP-graph replay of CG iteration chunks (small systems, profiling off): FP64 CG and
mixed FP32/FP64 CG on a 12^3 trilinear Laplace system with a lognormal
conductivity, solved with graph replay several times with eager work and
synchronisations between the solves, each result bit for bit equal to the solve
launched kernel by kernel."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _system():
    from iron_amd import basis, fem, mesh
    b = basis.basis(3, 1)
    m = mesh.RegularMesh((12, 12, 12), [b], (1.0, 1.0, 1.0))
    dm = mesh.DofMap(m, [0])
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, colours=mesh.structured_colours(m))
    X = m.coords
    fixed = (X[:, 0] == 0.0) | (X[:, 0] == 1.0)
    g = torch.Generator().manual_seed(4)
    kappa = torch.randn(m.num_elements, dtype=torch.float64, generator=g).exp().cuda()
    v = csr.zeros()
    fem.ScalarAssembler(b, X, m.elem_nodes[0], csr).assemble(v, kappa=kappa)
    red = fem.DirichletReduction(csr, fixed)
    S = fem.SolverMatrix(red)
    u = torch.as_tensor(np.where(X[:, 0] == 1.0, 1.0, 0.0), device="cuda")
    bvec = red.rhs(v, u)
    s64, s32 = S.gather2(v)
    return fem, red, S, s64, s32, bvec


def _solve(fem, red, S, s64, s32, bvec, ws, mixed):
    x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
    if mixed:
        info = fem.refine(S, s64, s32, bvec, x, 1e-10, 100000, ws)
    else:
        info = fem.cg(S, s64, bvec, x, 1e-10, 100000, ws)
    torch.cuda.synchronize()
    return x.cpu().numpy(), info


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_replays_equal_eager_launches(mixed):
    from iron_amd import _lib
    print(f"case CG graph replay mixed={mixed}", flush=True)
    fem, red, S, s64, s32, bvec = _system()
    ws = fem.KrylovWorkspace(red.n_free, "cuda")
    _lib.graphs_enable(False)
    try:
        ref, ref_info = _solve(fem, red, S, s64, s32, bvec, ws, mixed)
    finally:
        _lib.graphs_enable(True)
    assert ref_info.relres <= 1e-10
    junk = torch.rand(1 << 20, device="cuda")
    for rep in range(3):
        x, info = _solve(fem, red, S, s64, s32, bvec, ws, mixed)
        assert _lib.lib().mife_graphs_cached() > 0
        print(f"replay {rep}: {info.iterations} iterations, relres {info.relres:.2e}", flush=True)
        assert info.iterations == ref_info.iterations
        assert np.array_equal(x, ref)
        # eager work and a synchronisation between replays
        junk.mul_(1.5).add_(0.25)
        torch.cuda.synchronize()
        _lib.graphs_enable(False)
        try:
            x_e, _ = _solve(fem, red, S, s64, s32, bvec, ws, mixed)
        finally:
            _lib.graphs_enable(True)
        assert np.array_equal(x_e, ref)
