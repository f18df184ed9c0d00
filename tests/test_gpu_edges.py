"""Edge cases through the C ABI: zero-sized calls, a single element, systems
smaller than one 64-row slice, thin and ragged boxes, and a problem whose every
DOF is fixed.  Each case is checked against the FP64 oracle's dense solve."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402


def test_zero_sized_calls_leave_outputs_untouched():
    from iron_amd import _lib, basis, fem
    print("case zero-sized calls", flush=True)
    L = _lib.lib()
    y = torch.full((8,), 7.0, dtype=torch.float64, device="cuda")
    d = _lib.MatrixDesc(0, _lib.MIFE_FMT_CSR, y.data_ptr(), y.data_ptr(), None, 0, 0, None, None, 0)
    assert L.mife_spmv(C.byref(d), y.data_ptr(), y.data_ptr(), y.data_ptr(), 0, None) == 0
    db = fem.DeviceBasis(basis.basis(3, 1, 2), "cuda")
    p = y.data_ptr()
    rc = L.mife_assemble_scalar(C.byref(db.desc), 0, p, p, p, None, None, p, None, None, 0, 0, p, None, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.all(y == 7.0)


def _solve_box(ne, order, fixed_fn, ufix_fn, fmt, mixed=False):
    """GPU solve of a lognormal-kappa Poisson problem (f = 1) on a box; returns the
    GPU solution and the oracle's dense solution."""
    from iron_amd import basis, fem, mesh
    pb = basis.basis(3, order, order + 1)
    pm = mesh.RegularMesh(ne, [pb], (1.0, 1.0, 1.0))
    pd = mesh.DofMap(pm, [0])
    X = pm.coords
    fixed = fixed_fn(X)
    ufix = ufix_fn(X)
    rng = np.random.default_rng(sum(ne) + 10 * order)
    kappa = np.exp(rng.standard_normal(pm.num_elements))
    csr = fem.CSRStructure(pd.element_dofs(), pd.num_dofs)
    asm = fem.ScalarAssembler(pb, X, pm.elem_nodes[0], csr)
    v64 = csr.zeros(fem.F64)
    src = torch.zeros(pd.num_dofs, dtype=torch.float64, device="cuda")
    ones = torch.ones(pm.num_elements, dtype=torch.float64, device="cuda")
    asm.assemble(v64, kappa=torch.as_tensor(kappa, device="cuda"), source=ones, src_vec=src)
    red = fem.DirichletReduction(csr, fixed)
    u = torch.as_tensor(np.where(fixed, ufix, 0.0), device="cuda")
    if red.n_free > 0:
        S = fem.SolverMatrix(red, fmt)
        b = red.rhs(v64, u, src=src)
        x = torch.zeros(red.n_free, dtype=torch.float64, device="cuda")
        ws = fem.KrylovWorkspace(red.n_free, "cuda")
        if mixed:
            v32 = csr.zeros(fem.F32)
            asm.assemble(v32, kappa=torch.as_tensor(kappa, device="cuda"), prec=fem.F32)
            info = fem.refine(S, S.gather(v64), S.gather(v32, fem.F32), b, x, 1e-10, 100000, ws)
        else:
            info = fem.cg(S, S.gather(v64), b, x, 1e-12, 100000, ws)
        assert info.relres <= (1e-10 if mixed else 1e-12)
        red.field_update(x, u)
    # oracle: dense FP64 solve of the same reduced system
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0])
    xe = om.coords[om.elem_nodes[0] - 1]
    rp, col = fo.csr_pattern(od)
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    sref = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    uo = np.where(fixed, ufix, 0.0)
    if (~fixed).any():
        free, srp, scol, sv, bref = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, sref)
        uo[free] = np.linalg.solve(fo._csr_scipy(srp, scol, sv).toarray(), bref)
    return u.cpu().numpy(), uo


def _x_faces(X):
    return (X[:, 0] == 0.0) | (X[:, 0] == 1.0)


def _x_ramp(X):
    return np.where(X[:, 0] == 1.0, 1.0, 0.0)


@pytest.mark.parametrize("ne,order,fmt", [((1, 1, 1), 1, "sell"), ((1, 1, 1), 2, "sell"), ((1, 1, 1), 2, "csr"),
                                          ((2, 1, 1), 2, "sell"), ((7, 1, 1), 1, "sell"), ((1, 9, 2), 1, "sell"),
                                          ((3, 5, 1), 2, "sell"), ((3, 5, 1), 2, "csr")])
def test_tiny_thin_and_ragged_boxes(ne, order, fmt):
    """Fewer free rows than one slice (1x1x1), one-element-thick boxes and odd
    extents, in both solver storages, against the dense oracle solve."""
    print(f"case box {ne} Q{order} {fmt}", flush=True)
    u, uo = _solve_box(ne, order, _x_faces, _x_ramp, fmt)
    assert np.abs(u - uo).max() <= 1e-10 * max(1.0, np.abs(uo).max())


def test_single_free_dof_and_mixed_precision():
    """A 2x1x1 trilinear box fixed everywhere except one node, and the mixed FP32/FP64
    path on a thin box."""
    print("case single free DOF", flush=True)

    def all_but_one(X):
        f = np.ones(len(X), dtype=bool)
        f[np.argmin(np.abs(X[:, 0] - 0.5) + X[:, 1] + X[:, 2])] = False
        return f

    u, uo = _solve_box((2, 1, 1), 1, all_but_one, _x_ramp, "sell")
    assert np.abs(u - uo).max() <= 1e-12 * max(1.0, np.abs(uo).max())
    u, uo = _solve_box((6, 1, 3), 2, _x_faces, _x_ramp, "sell", mixed=True)
    assert np.abs(u - uo).max() <= 1e-8 * max(1.0, np.abs(uo).max())


def test_every_dof_fixed_through_the_api():
    """Every node carries a fixed value: no solve runs and the field equals the BCs."""
    from iron_amd import iron
    print("case every DOF fixed (API)", flush=True)
    cs = iron.CoordinateSystem()
    cs.CreateStart(1)
    cs.dimension = 3
    cs.CreateFinish()
    region = iron.Region()
    region.CreateStart(1, iron.WorldRegion)
    region.coordinateSystem = cs
    region.CreateFinish()
    b = iron.Basis()
    b.CreateStart(1)
    b.type = iron.BasisTypes.LAGRANGE_HERMITE_TP
    b.numberOfXi = 3
    b.interpolationXi = [iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE] * 3
    b.CreateFinish()
    gm = iron.GeneratedMesh()
    gm.CreateStart(1, region)
    gm.type = iron.GeneratedMeshTypes.REGULAR
    gm.basis = [b]
    gm.extent = [1.0, 1.0, 1.0]
    gm.numberOfElements = [2, 2, 1]
    mesh = iron.Mesh()
    gm.CreateFinish(1, mesh)
    dec = iron.Decomposition()
    dec.CreateStart(1, mesh)
    dec.CreateFinish()
    geo = iron.Field()
    geo.CreateStart(1, region)
    geo.meshDecomposition = dec
    geo.CreateFinish()
    gm.GeometricParametersCalculate(geo)
    es = iron.EquationsSet()
    esf = iron.Field()
    es.CreateStart(1, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD, iron.EquationsSetTypes.LAPLACE_EQUATION,
                                    iron.EquationsSetSubtypes.STANDARD_LAPLACE], 2, esf)
    es.CreateFinish()
    dep = iron.Field()
    es.DependentCreateStart(3, dep)
    es.DependentCreateFinish()
    eq = iron.Equations()
    es.EquationsCreateStart(eq)
    es.EquationsCreateFinish()
    problem = iron.Problem()
    problem.CreateStart(1, [iron.ProblemClasses.CLASSICAL_FIELD, iron.ProblemTypes.LAPLACE_EQUATION,
                            iron.ProblemSubtypes.STANDARD_LAPLACE])
    problem.CreateFinish()
    problem.ControlLoopCreateStart()
    problem.ControlLoopCreateFinish()
    solver = iron.Solver()
    problem.SolversCreateStart()
    problem.SolverGet([iron.ControlLoopIdentifiers.NODE], 1, solver)
    problem.SolversCreateFinish()
    se = iron.SolverEquations()
    problem.SolverEquationsCreateStart()
    solver.SolverEquationsGet(se)
    se.EquationsSetAdd(es)
    problem.SolverEquationsCreateFinish()
    bc = iron.BoundaryConditions()
    se.BoundaryConditionsCreateStart(bc)
    n_nodes = 3 * 3 * 2
    for node in range(1, n_nodes + 1):
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, node, 1, iron.BoundaryConditionsTypes.FIXED, 0.1 * node)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, iron.FieldParameterSetTypes.VALUES).copy()
    assert np.array_equal(u, 0.1 * np.arange(1, n_nodes + 1))
    assert solver.last_info.iterations == 0
