"""The iron-compatible API end to end on the GPU, against the FP64 oracle.

The scripts below are written the way the reference's Python examples are
(bindings/python/tests/Laplace.py), with ``from iron_amd import iron``.
"""
import json
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"


def _setup(iron, ne, interp, extent, origin=None, dim=3, gauss=None):
    cs = iron.CoordinateSystem()
    cs.CreateStart(1)
    cs.dimension = dim
    cs.CreateFinish()
    region = iron.Region()
    region.CreateStart(1, iron.WorldRegion)
    region.label = "Region"
    region.coordinateSystem = cs
    region.CreateFinish()
    basis = iron.Basis()
    basis.CreateStart(1)
    basis.type = iron.BasisTypes.LAGRANGE_HERMITE_TP
    basis.numberOfXi = dim
    basis.interpolationXi = [interp] * dim
    if gauss:
        basis.quadratureNumberOfGaussXi = [gauss] * dim
    basis.CreateFinish()
    gm = iron.GeneratedMesh()
    gm.CreateStart(1, region)
    gm.type = iron.GeneratedMeshTypes.REGULAR
    gm.basis = [basis]
    gm.extent = list(extent)
    gm.numberOfElements = list(ne)
    if origin is not None:
        gm.origin = list(origin)
    mesh = iron.Mesh()
    gm.CreateFinish(1, mesh)
    dec = iron.Decomposition()
    dec.CreateStart(1, mesh)
    dec.type = iron.DecompositionTypes.CALCULATED
    dec.numberOfDomains = iron.ComputationalNumberOfNodesGet()
    dec.CreateFinish()
    geo = iron.Field()
    geo.CreateStart(1, region)
    geo.meshDecomposition = dec
    for c in range(dim):
        geo.ComponentMeshComponentSet(iron.FieldVariableTypes.U, c + 1, 1)
    geo.CreateFinish()
    gm.GeometricParametersCalculate(geo)
    return region, dec, geo, mesh


def _problem(iron, es, spec, solver_cfg):
    problem = iron.Problem()
    problem.CreateStart(1, spec)
    problem.CreateFinish()
    problem.ControlLoopCreateStart()
    problem.ControlLoopCreateFinish()
    solver = iron.Solver()
    problem.SolversCreateStart()
    problem.SolverGet([iron.ControlLoopIdentifiers.NODE], 1, solver)
    for k, v in solver_cfg.items():
        setattr(solver, k, v)
    problem.SolversCreateFinish()
    solver = iron.Solver()
    se = iron.SolverEquations()
    problem.SolverEquationsCreateStart()
    problem.SolverGet([iron.ControlLoopIdentifiers.NODE], 1, solver)
    solver.SolverEquationsGet(se)
    se.sparsityType = iron.SolverEquationsSparsityTypes.SPARSE
    se.EquationsSetAdd(es)
    problem.SolverEquationsCreateFinish()
    bc = iron.BoundaryConditions()
    se.BoundaryConditionsCreateStart(bc)
    return problem, solver, se, bc


def _equations_set(iron, region, geo, spec):
    es = iron.EquationsSet()
    esf = iron.Field()
    es.CreateStart(1, region, geo, spec, 2, esf)
    es.CreateFinish()
    dep = iron.Field()
    es.DependentCreateStart(3, dep)
    dep.DOFOrderTypeSet(iron.FieldVariableTypes.U, iron.FieldDOFOrderTypes.SEPARATED)
    dep.DOFOrderTypeSet(iron.FieldVariableTypes.DELUDELN, iron.FieldDOFOrderTypes.SEPARATED)
    es.DependentCreateFinish()
    return es, dep


def _equations(iron, es):
    eq = iron.Equations()
    es.EquationsCreateStart(eq)
    eq.sparsityType = iron.EquationsSparsityTypes.SPARSE
    eq.outputType = iron.EquationsOutputTypes.NONE
    es.EquationsCreateFinish()


def _oracle_scalar(ne, order, extent, origin, kappa, source, fixed, ufix):
    ob = fo.Basis(nxi=len(ne), order=order, ngauss_xi=order + 1)
    om = fo.regular_mesh(ne, [ob], extent, origin)
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = np.zeros(od.num_dofs) if source is None else \
        fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, source), od.num_dofs)
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, src)
    A = fo._csr_scipy(srp, scol, sv)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(A.tocsc(), b)
    u = fo.field_update(od.num_dofs, free, x, fixed, ufix)
    return u, (rp, col, vals, src), (A, b, free)


def test_laplace_example_script():
    """bindings/python/tests/Laplace.py: 5x5x5 trilinear on 2x1x3, u(node 1)=0,
    u(last node)=1, GMRES defaults with 1e-12 tolerances, initial field 0.5."""
    from iron_amd import iron
    ne, extent = (5, 5, 5), (2.0, 1.0, 3.0)
    region, dec, geo, mesh = _setup(iron, ne, iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE, extent)
    es, dep = _equations_set(iron, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD,
                                                 iron.EquationsSetTypes.LAPLACE_EQUATION,
                                                 iron.EquationsSetSubtypes.STANDARD_LAPLACE])
    dep.ComponentValuesInitialiseDP(iron.FieldVariableTypes.U, iron.FieldParameterSetTypes.VALUES, 1, 0.5)
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.CLASSICAL_FIELD,
                                                  iron.ProblemTypes.LAPLACE_EQUATION,
                                                  iron.ProblemSubtypes.STANDARD_LAPLACE],
                                       {"outputType": iron.SolverOutputTypes.SOLVER,
                                        "linearType": iron.LinearSolverTypes.ITERATIVE,
                                        "linearIterativeAbsoluteTolerance": 1.0e-12,
                                        "linearIterativeRelativeTolerance": 1.0e-12})
    nodes = iron.Nodes()
    region.NodesGet(nodes)
    last = nodes.numberOfNodes
    assert last == 6 * 6 * 6
    bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, 1, 1, iron.BoundaryConditionsTypes.FIXED, 0.0)
    bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, last, 1, iron.BoundaryConditionsTypes.FIXED, 1.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, iron.FieldParameterSetTypes.VALUES).copy()
    fixed = np.zeros(last, bool)
    fixed[[0, last - 1]] = True
    ufix = np.zeros(last)
    ufix[last - 1] = 1.0
    uo, _, _ = _oracle_scalar(ne, 1, extent, None, None, None, fixed, ufix)
    assert np.max(np.abs(u - uo)) < 1e-9
    assert solver.last_info.relres <= 1e-11
    assert dep.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, last, 1) == 1.0


def test_analytic_laplace_golden_through_api():
    """tests/ClassicalField/AnalyticLaplace.F90 export case through the API: the
    reference's .opanal nodal values and 'del Phi/del n' (iron-compatible)."""
    from iron_amd import iron
    data = json.loads((GOLDEN / "analytic_laplace_linear_lagrange.json").read_text())
    o = -data["setup"]["origin_pi"] / 2.0
    region, dec, geo, mesh = _setup(iron, (2, 2), iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE,
                                    (2.0, 2.0), origin=(o, o), dim=2)
    es, dep = _equations_set(iron, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD,
                                                 iron.EquationsSetTypes.LAPLACE_EQUATION,
                                                 iron.EquationsSetSubtypes.STANDARD_LAPLACE])
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.CLASSICAL_FIELD,
                                                  iron.ProblemTypes.LAPLACE_EQUATION,
                                                  iron.ProblemSubtypes.STANDARD_LAPLACE],
                                       {"linearIterativeAbsoluteTolerance": 1.0e-12,
                                        "linearIterativeRelativeTolerance": 1.0e-12})
    for node in range(1, 10):
        if node == 5:
            continue
        x = geo.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, node, 1)
        y = geo.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, node, 2)
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, node, 1, iron.BoundaryConditionsTypes.FIXED,
                   np.cos(x) * np.cosh(y))
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    for node, (num, _) in data["nodal"]["Phi"].items():
        v = dep.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, int(node), 1)
        assert float(f"{v:.4e}") == pytest.approx(num, rel=1e-4, abs=1e-15)
    for node, (num, _) in data["nodal"]["del Phi/del n"].items():
        v = dep.ParameterSetGetNodeDP(iron.FieldVariableTypes.DELUDELN, 1, 1, 1, int(node), 1)
        if abs(num) < 1e-10:
            assert abs(v) < 1e-12
        else:
            assert v == pytest.approx(num, rel=1e-4)


def test_generalised_laplace_with_source_mixed_precision():
    """Element conductivity + Poisson source, FP32 inner CG with FP64 residual
    replacement: 1e-8 relative residual in the oracle's FP64 system."""
    from iron_amd import iron
    ne = (5, 4, 4)
    rng = np.random.default_rng(4)
    kappa = np.exp(rng.standard_normal(int(np.prod(ne))))
    region, dec, geo, mesh = _setup(iron, ne, iron.BasisInterpolationSpecifications.QUADRATIC_LAGRANGE,
                                    (1.0, 1.0, 1.0))
    es, dep = _equations_set(iron, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD,
                                                 iron.EquationsSetTypes.LAPLACE_EQUATION,
                                                 iron.EquationsSetSubtypes.GENERALISED_LAPLACE])
    mat = iron.Field()
    es.MaterialsCreateStart(4, mat)
    for c in (1, 2, 3):
        mat.ComponentInterpolationSet(iron.FieldVariableTypes.U, c, iron.FieldInterpolationTypes.ELEMENT_BASED)
    es.MaterialsCreateFinish()
    for c in (1, 2, 3):
        for e in range(len(kappa)):
            mat.ParameterSetUpdateElementDP(iron.FieldVariableTypes.U, 1, e + 1, c, kappa[e])
    src = iron.Field()
    es.SourceCreateStart(5, src)
    es.SourceCreateFinish()
    src.ComponentValuesInitialiseDP(iron.FieldVariableTypes.U, 1, 1, 1.0)
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.CLASSICAL_FIELD,
                                                  iron.ProblemTypes.LAPLACE_EQUATION,
                                                  iron.ProblemSubtypes.GENERALISED_LAPLACE],
                                       {"linearIterativeType": iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT,
                                        "linearIterativeRelativeTolerance": 1.0e-8,
                                        "linearIterativeAbsoluteTolerance": 0.0,
                                        "linearIterativePrecision": "mixed"})
    X = np.stack([geo._component(iron.FieldVariableTypes.U, c) for c in (1, 2, 3)], axis=1)
    fixed = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1.0) < 1e-12), axis=1)
    rm = geo._mesh()
    for i in np.nonzero(fixed)[0]:
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(rm.comp_user[0][i]), 1,
                   iron.BoundaryConditionsTypes.FIXED, 0.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    uo, (rp, col, vals, srcv), (A, b, free) = _oracle_scalar(ne, 2, (1.0, 1.0, 1.0), None, kappa, 1.0, fixed,
                                                             np.zeros(len(u)))
    rel = np.linalg.norm(b - A @ u[free]) / np.linalg.norm(b)
    assert rel <= 1e-8
    assert np.max(np.abs(u - uo)) <= 1e-7 * np.max(np.abs(uo))
    assert solver.last_info.outer >= 2
    # correct reactions: K u - f, zero on free rows
    R = es.reactions
    assert np.max(np.abs(R[free])) < 1e-7 * np.max(np.abs(R))


def test_linear_elasticity_cantilever_traction():
    """3-D linear elasticity, clamped at x=0, NEUMANN_POINT traction on x=3 with
    NEUMANN_INTEGRATED_ONLY on the adjacent boundary ring, element-based E."""
    from iron_amd import iron
    ne, extent = (6, 2, 2), (3.0, 1.0, 1.0)
    rng = np.random.default_rng(9)
    E = 10.0 * np.exp(0.3 * rng.standard_normal(int(np.prod(ne))))
    region, dec, geo, mesh = _setup(iron, ne, iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE, extent)
    es, dep = _equations_set(iron, region, geo, [iron.EquationsSetClasses.ELASTICITY,
                                                 iron.EquationsSetTypes.LINEAR_ELASTICITY,
                                                 iron.EquationsSetSubtypes.THREE_DIMENSIONAL])
    mat = iron.Field()
    es.MaterialsCreateStart(4, mat)
    for c in (1, 2, 3):
        mat.ComponentInterpolationSet(iron.FieldVariableTypes.U, c, iron.FieldInterpolationTypes.ELEMENT_BASED)
    es.MaterialsCreateFinish()
    for c in (1, 2, 3):
        mat.ParameterSetDataGet(iron.FieldVariableTypes.U, 1)[mat._offset(iron.FieldVariableTypes.U, c):
                                                              mat._offset(iron.FieldVariableTypes.U, c) + len(E)] = E
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.ELASTICITY,
                                                  iron.ProblemTypes.LINEAR_ELASTICITY,
                                                  iron.ProblemSubtypes.NONE],
                                       {"linearIterativeType": iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT,
                                        "linearIterativeRelativeTolerance": 1.0e-10,
                                        "linearIterativeAbsoluteTolerance": 0.0})
    rm = geo._mesh()
    X = rm.coords
    users = rm.comp_user[0]
    clamp = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
    face = np.nonzero(np.abs(X[:, 0] - 3.0) < 1e-12)[0]
    hx = 3.0 / ne[0]
    onb = (np.abs(X[:, 1]) < 1e-12) | (np.abs(X[:, 1] - 1) < 1e-12) | (np.abs(X[:, 2]) < 1e-12) | \
          (np.abs(X[:, 2] - 1) < 1e-12)
    ring = np.nonzero((np.abs(X[:, 0] - (3.0 - hx)) < 1e-12) & onb)[0]
    for i in clamp:
        for c in (1, 2, 3):
            bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), c, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in face:
        bc.SetNode(dep, iron.FieldVariableTypes.DELUDELN, 1, 1, int(users[i]), 3,
                   iron.BoundaryConditionsTypes.NEUMANN_POINT, -1.0)
    for i in ring:
        bc.SetNode(dep, iron.FieldVariableTypes.DELUDELN, 1, 1, int(users[i]), 3,
                   iron.BoundaryConditionsTypes.NEUMANN_INTEGRATED_ONLY, 0.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    # whole nodes clamped: the engine stores the solver matrix as 3x3 node blocks
    assert es._device_cache.S.fmt == "sell3"
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    # oracle pipeline
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    nu = 0.3 * np.ones_like(E)
    C = fo.linear_elasticity_tensor(E, E, E, nu, nu, nu)
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs)
    n = od.num_dofs
    N1 = len(X)
    fixed = np.zeros(n, bool)
    for c in range(3):
        fixed[c * N1 + clamp] = True
    pv = {int(2 * N1 + i): -1.0 for i in face}
    io = frozenset(int(2 * N1 + i) for i in ring)
    rhs_fixed = np.zeros(n, bool)
    rhs_fixed[list(pv)] = True
    rhs_fixed[list(io)] = True
    f = fo.neumann_integrate(om, od, pv, io)
    assert abs(f.sum() + 1.0) < 1e-12  # uniform unit traction on a unit face
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, np.zeros(n), f, rhs_fixed)
    A = fo._csr_scipy(srp, scol, sv)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(A.tocsc(), b)
    assert np.max(np.abs(u[free] - x)) < 1e-8 * np.max(np.abs(x))
    assert np.all(u[fixed] == 0.0)
    # tip deflects in -z
    assert u[2 * N1 + face].mean() < 0


def test_krylov_iteration_limit_warns_and_keeps_the_iterate():
    """A KSP that stops at its iteration limit is a warning in iron, not an error
    (PETSC_KSP_DIVERGED_ITS -> FLAG_WARNING, solver_routines.F90:11875-11880): the
    field still takes the last iterate."""
    from iron_amd import iron
    print("KSP iteration limit through the API", flush=True)
    ne = (6, 5, 4)
    region, dec, geo, mesh = _setup(iron, ne, iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE, (1.0, 1.0, 1.0))
    es, dep = _equations_set(iron, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD,
                                                 iron.EquationsSetTypes.LAPLACE_EQUATION,
                                                 iron.EquationsSetSubtypes.STANDARD_LAPLACE])
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.CLASSICAL_FIELD,
                                                  iron.ProblemTypes.LAPLACE_EQUATION,
                                                  iron.ProblemSubtypes.STANDARD_LAPLACE],
                                       {"linearIterativeType": iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT,
                                        "linearIterativeRelativeTolerance": 1.0e-12,
                                        "linearIterativeAbsoluteTolerance": 0.0,
                                        "linearIterativeMaximumIterations": 2})
    rm = geo._mesh()
    X = rm.coords
    users = rm.comp_user[0]
    for i in np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]:
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), 1, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in np.nonzero(np.abs(X[:, 0] - 1) < 1e-12)[0]:
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), 1, iron.BoundaryConditionsTypes.FIXED, 1.0)
    se.BoundaryConditionsCreateFinish()
    with pytest.warns(RuntimeWarning, match="did not converge"):
        problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    interior = (X[:, 0] > 1e-12) & (X[:, 0] < 1 - 1e-12)
    assert np.all(np.isfinite(u)) and np.any(u[interior] != 0.0)
    assert np.max(np.abs(u[interior] - X[interior, 0])) > 1e-6  # two iterations are not the solution
    assert solver.last_info.iterations == 2
