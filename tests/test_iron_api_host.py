"""Host-side behaviour of the iron-compatible API (no GPU needed)."""
import numpy as np
import pytest

from iron_amd import iron


def test_enum_values_match_iron_constants():
    # values from src/*_routines.F90 / *_constants.F90 aliased by src/opencmiss_iron.F90
    assert iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT == 3
    assert iron.IterativeLinearSolverTypes.GMRES == 5
    assert iron.IterativePreconditionerTypes.JACOBI == 1
    assert iron.LinearSolverTypes.ITERATIVE == 2
    assert iron.EquationsSetClasses.CLASSICAL_FIELD == 4 and iron.EquationsSetClasses.ELASTICITY == 1
    assert iron.EquationsSetSubtypes.GENERALISED_LAPLACE == 2
    assert iron.EquationsSetSubtypes.THREE_DIMENSIONAL == 4
    assert iron.FieldVariableTypes.DELUDELN == 2 and iron.FieldVariableTypes.V == 5
    assert iron.FieldInterpolationTypes.ELEMENT_BASED == 2
    assert iron.BoundaryConditionsTypes.NEUMANN_POINT == 8
    assert iron.BoundaryConditionsTypes.FIXED_INCREMENTED == 13
    assert iron.BoundaryConditionsTypes.NEUMANN_INTEGRATED_ONLY == 20


def test_set_get_method_forms_alias_properties():
    cs = iron.CoordinateSystem()
    cs.CreateStart(1)
    cs.DimensionSet(2)
    assert cs.dimension == 2 and cs.DimensionGet() == 2
    s = iron.Solver()
    s.LinearIterativeRelativeToleranceSet(1e-9)
    assert s.linearIterativeRelativeTolerance == 1e-9
    # iron's KSP defaults (solver_routines.F90:11320-11369)
    assert s.linearIterativeType == iron.IterativeLinearSolverTypes.GMRES
    assert s.linearIterativeMaximumIterations == 100000


def test_unsupported_specification_raises_cmfe_error():
    es = iron.EquationsSet()
    with pytest.raises(iron.CMFEError):
        es.CreateStart(1, None, None, [iron.EquationsSetClasses.CLASSICAL_FIELD, 7, 1], 2, None)


def test_field_parameter_access_and_mesh_numbering():
    cs = iron.CoordinateSystem(); cs.CreateStart(1); cs.CreateFinish()
    region = iron.Region(); region.CreateStart(1, iron.WorldRegion); region.CreateFinish()
    basis = iron.Basis(); basis.CreateStart(1)
    basis.interpolationXi = [iron.BasisInterpolationSpecifications.QUADRATIC_LAGRANGE] * 3
    basis.CreateFinish()
    gm = iron.GeneratedMesh(); gm.CreateStart(1, region)
    gm.basis = [basis]; gm.extent = [2.0, 1.0, 1.0]; gm.numberOfElements = [2, 1, 1]
    mesh = iron.Mesh(); gm.CreateFinish(1, mesh)
    dec = iron.Decomposition(); dec.CreateStart(1, mesh); dec.CreateFinish()
    geo = iron.Field(); geo.CreateStart(1, region); geo.meshDecomposition = dec; geo.CreateFinish()
    gm.GeometricParametersCalculate(geo)
    nodes = iron.Nodes(); region.NodesGet(nodes)
    assert nodes.numberOfNodes == 5 * 3 * 3
    # node 2 is the mid-edge node at x = 0.5
    assert geo.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, 2, 1) == 0.5
    assert geo.ParameterSetGetNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, 45, 3) == 1.0


def _laplace_set(ne=(3, 2, 2)):
    cs = iron.CoordinateSystem(); cs.CreateStart(1); cs.CreateFinish()
    region = iron.Region(); region.CreateStart(1, iron.WorldRegion); region.CreateFinish()
    basis = iron.Basis(); basis.CreateStart(1)
    basis.interpolationXi = [iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE] * 3
    basis.CreateFinish()
    gm = iron.GeneratedMesh(); gm.CreateStart(1, region)
    gm.basis = [basis]; gm.extent = [1.0, 1.0, 1.0]; gm.numberOfElements = list(ne)
    mesh = iron.Mesh(); gm.CreateFinish(1, mesh)
    dec = iron.Decomposition(); dec.CreateStart(1, mesh); dec.CreateFinish()
    geo = iron.Field(); geo.CreateStart(1, region); geo.meshDecomposition = dec; geo.CreateFinish()
    gm.GeometricParametersCalculate(geo)
    es = iron.EquationsSet()
    es.CreateStart(1, region, geo, [iron.EquationsSetClasses.CLASSICAL_FIELD, iron.EquationsSetTypes.LAPLACE_EQUATION,
                                    iron.EquationsSetSubtypes.STANDARD_LAPLACE], 2, iron.Field())
    es.CreateFinish()
    dep = iron.Field()
    es.DependentCreateStart(3, dep)
    es.DependentCreateFinish()
    return es, dep, geo


def test_dof_constraint_mapping_and_the_reference_error_rules():
    """ConstrainNodeDofsEqual -> solver DOF and coefficient per DOF, and iron's checks
    (DofConstraintSet / DofConstraintsCreateFinish, boundary_condition_routines.F90:3204-3472)."""
    from iron_amd import engine
    es, dep, geo = _laplace_set()
    rm = geo._mesh()
    dm = engine._dof_map(dep, rm)
    n = dm.num_dofs
    U = iron.FieldVariableTypes.U
    fixed = np.zeros(n, bool)
    fixed[0] = True  # node 1
    bc = iron.BoundaryConditions()
    bc.ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [2, 5, 9], 0.5)
    owner, coef, constrained = engine._dof_constraints(bc, es, dm, rm, fixed)
    assert constrained.sum() == 2 and constrained[4] and constrained[8]
    assert owner[0] == -1 and owner[4] == owner[1] == owner[8] == 0 and owner[2] == 1
    assert coef[4] == coef[8] == 0.5 and coef[1] == 1.0
    assert owner.max() == n - 1 - 2 - 1
    with pytest.raises(iron.CMFEError):  # the DOF depended on is fixed
        b2 = iron.BoundaryConditions(); b2.ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [1, 3], 1.0)
        engine._dof_constraints(b2, es, dm, rm, fixed)
    with pytest.raises(iron.CMFEError):  # a constrained DOF that is also fixed
        b3 = iron.BoundaryConditions(); b3.ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [3, 1], 1.0)
        engine._dof_constraints(b3, es, dm, rm, fixed)
    with pytest.raises(iron.CMFEError):  # constrained twice
        b4 = iron.BoundaryConditions()
        b4.ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [2, 3], 1.0)
        b4.ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [4, 3], 1.0)
        engine._dof_constraints(b4, es, dm, rm, fixed)
    with pytest.raises(iron.CMFEError):  # fewer than two nodes
        iron.BoundaryConditions().ConstrainNodeDofsEqual(dep, U, 1, 1, 1, [2], 1.0)
    assert engine._dof_constraints(iron.BoundaryConditions(), es, dm, rm, fixed) is None


def _mr_set():
    cs = iron.CoordinateSystem(); cs.CreateStart(1); cs.CreateFinish()
    region = iron.Region(); region.CreateStart(1, iron.WorldRegion); region.CreateFinish()
    basis = iron.Basis(); basis.CreateStart(1)
    basis.interpolationXi = [iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE] * 3
    basis.CreateFinish()
    gm = iron.GeneratedMesh(); gm.CreateStart(1, region)
    gm.basis = [basis]; gm.extent = [1.0, 1.0, 1.0]; gm.numberOfElements = [2, 1, 1]
    mesh = iron.Mesh(); gm.CreateFinish(1, mesh)
    dec = iron.Decomposition(); dec.CreateStart(1, mesh); dec.CreateFinish()
    geo = iron.Field(); geo.CreateStart(1, region); geo.meshDecomposition = dec; geo.CreateFinish()
    gm.GeometricParametersCalculate(geo)
    es = iron.EquationsSet()
    es.CreateStart(1, region, geo, [iron.EquationsSetClasses.ELASTICITY, iron.EquationsSetTypes.FINITE_ELASTICITY,
                                    iron.EquationsSetSubtypes.MOONEY_RIVLIN], 2, iron.Field())
    es.CreateFinish()
    dep = iron.Field()
    es.DependentCreateStart(3, dep)
    dep.ComponentInterpolationSet(iron.FieldVariableTypes.U, 4, iron.FieldInterpolationTypes.ELEMENT_BASED)
    es.DependentCreateFinish()
    return es, dep, geo


def test_finite_elasticity_rejects_conditions_on_another_field():
    """A condition set on the geometric field is an error (as on the linear path),
    raised before any device work."""
    from iron_amd import finite_elasticity
    es, dep, geo = _mr_set()
    bc = iron.BoundaryConditions()
    bc.SetNode(geo, iron.FieldVariableTypes.U, 1, 1, 1, 1, iron.BoundaryConditionsTypes.FIXED, 0.0)
    with pytest.raises(iron.CMFEError, match="dependent field"):
        finite_elasticity.solve(None, es, None, bc)


def test_geometry_is_read_from_the_geometric_field():
    """Element kernels, Neumann integration and finite elasticity take node
    coordinates from the geometric field's values: moved nodes are seen, and a
    geometric field not node based on the first mesh component is refused."""
    from iron_amd import engine
    es, dep, geo = _laplace_set()
    rm = geo._mesh()
    X = engine.geometry_coords(geo, rm, 0)
    assert np.array_equal(X, rm.coords)
    geo.ParameterSetUpdateNodeDP(iron.FieldVariableTypes.U, 1, 1, 1, 6, 2, 0.4)
    X2 = engine.geometry_coords(geo, rm, 0)
    k = int(rm.local_node(6, 0))
    assert X2[k, 1] == 0.4 and np.array_equal(np.delete(X2, k, 0), np.delete(rm.coords, k, 0))
    with pytest.raises(iron.CMFEError):
        engine.geometry_coords(geo, rm, 1)


def _q2_mesh_with_linear_component(ne=(2, 3, 2), dims=3):
    cs = iron.CoordinateSystem(); cs.CreateStart(1); cs.CreateFinish()
    region = iron.Region(); region.CreateStart(1, iron.WorldRegion); region.CreateFinish()
    bases = []
    for k, interp in enumerate((iron.BasisInterpolationSpecifications.QUADRATIC_LAGRANGE,
                                iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE)):
        b = iron.Basis(); b.CreateStart(k + 1); b.numberOfXi = dims
        b.interpolationXi = [interp] * dims
        b.CreateFinish()
        bases.append(b)
    gm = iron.GeneratedMesh(); gm.CreateStart(1, region)
    gm.basis = bases; gm.extent = [2.0, 1.0, 1.0][:dims]; gm.numberOfElements = list(ne)[:dims]
    mesh = iron.Mesh(); gm.CreateFinish(1, mesh)
    return gm, mesh._regular


def test_regular_surface_nodes_and_normals():
    """GENERATED_MESH_REGULAR_SURFACE_GET (generated_mesh_routines.F90:3943-4084): the
    face's nodes in the reference's loop order (LEFT/RIGHT: k outer, j inner; TOP/BOTTOM:
    j outer, i inner; FRONT/BACK: k outer, i inner), rebuilt here from node coordinates,
    for the quadratic component and the linear (corner-node) component."""
    gm, rm = _q2_mesh_with_linear_component()
    S = iron.GeneratedMeshSurfaceTypes
    X = rm.coords
    user0 = np.arange(1, len(X) + 1)
    expect = {S.REGULAR_LEFT: (0, 0.0, (1, 2), -1), S.REGULAR_RIGHT: (0, 2.0, (1, 2), 1),
              S.REGULAR_FRONT: (1, 0.0, (0, 2), -2), S.REGULAR_BACK: (1, 1.0, (0, 2), 2),
              S.REGULAR_BOTTOM: (2, 0.0, (0, 1), -3), S.REGULAR_TOP: (2, 1.0, (0, 1), 3)}
    corner = np.zeros(len(X), bool)
    corner[rm.comp_user[1] - 1] = True
    for st, (axis, val, (inner, outer), normal) in expect.items():
        for mc, sel in ((1, np.ones(len(X), bool)), (2, corner)):
            on = sel & (X[:, axis] == val)
            order = np.lexsort((X[on, inner], X[on, outer]))
            nodes, nx = gm.SurfaceGet(mc, st) if mc == 2 else gm.SurfaceGet(st)
            assert nx == normal
            assert np.array_equal(nodes, user0[on][order]), (st, mc)
    n_top, _ = gm.SurfaceGet(S.REGULAR_TOP)
    assert len(n_top) == 5 * 7
    with pytest.raises(iron.CMFEError):
        gm.SurfaceGet(7)
    # a 2-D mesh has one layer in xi3: TOP and BOTTOM are every node of the component
    gm2, rm2 = _q2_mesh_with_linear_component((2, 2), dims=2)
    assert np.array_equal(gm2.SurfaceGet(S.REGULAR_TOP)[0], rm2.comp_user[0])
    assert np.array_equal(gm2.SurfaceGet(2, S.REGULAR_BOTTOM)[0], rm2.comp_user[1])


def test_nearly_incompressible_subtype_fields():
    """EQUATIONS_SET_NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE (equations_set_constants.F90:146):
    a 3-component dependent field (no pressure) and materials c1, c2, kappa."""
    assert iron.EquationsSetSubtypes.NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN == 22
    gm, rm = _q2_mesh_with_linear_component((1, 1, 1))
    es = iron.EquationsSet()
    es.CreateStart(1, None, None, [iron.EquationsSetClasses.ELASTICITY, iron.EquationsSetTypes.FINITE_ELASTICITY,
                                   iron.EquationsSetSubtypes.NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN], 2, None)
    dec = iron.Decomposition(); dec.CreateStart(1, gm._mesh); dec.CreateFinish()
    geo = iron.Field(); geo.meshDecomposition = dec
    es.geometricField = geo
    dep, mat = iron.Field(), iron.Field()
    es.DependentCreateStart(3, dep)
    es.MaterialsCreateStart(4, mat)
    assert dep._var(iron.FieldVariableTypes.U).ncomp == 3
    assert mat._var(iron.FieldVariableTypes.U).ncomp == 3
