"""Linear DOF constraints on the GPU (mife_couple_* through the C ABI), against the
oracle's coupled solver mapping (fe_oracle.constraint_couplings / reduce_coupled),
which tests/test_constraints_oracle.py pins to the KKT form, P^T K P and the
closed form u = x.

Tolerances: the solver pattern and the row groups are integer data and must be
identical.  Solver values sum the same products in the same order as the oracle;
the device may contract a multiply into the add, so they are compared to 1e-14 of
the largest entry.  FP32 values to 1e-6 relative.  Solutions to the solver
tolerance stated in each test.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402
from test_constraints_oracle import _face_ties  # noqa: E402
from test_gpu_iron_api import _equations, _equations_set, _problem, _setup  # noqa: E402


def _dense(rp, col, vals, n):
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(rp)), col] = vals
    return A


@pytest.mark.parametrize("order", [1, 2])
def test_coupled_pattern_values_rhs_and_solve_match_the_oracle(order):
    from iron_amd import fem
    ne = (4, 3, 3)
    print(f"coupled reduction order={order} ne={ne}", flush=True)
    rng = np.random.default_rng(11 + order)
    ob = fo.Basis(nxi=3, order=order, ngauss_xi=order + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    kappa = np.exp(rng.standard_normal(om.num_elements))
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, 1.0), od.num_dofs)
    X = om.coords
    n = od.num_dofs
    fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1) < 1e-12)
    ufix = np.where(np.abs(X[:, 0]) < 1e-12, 0.25, 1.0)
    cf = {}
    ties = _face_ties(X, lambda i: cf.setdefault(i, float(rng.uniform(-1.5, 1.5))))
    owner, coef = fo.constraint_couplings(fixed, ties)
    constrained = np.zeros(n, bool)
    constrained[[c for c, _, _ in ties]] = True
    srp, scol, sv, b = fo.reduce_coupled(rp, col, vals, owner, coef, ufix, src)
    n_s = len(srp) - 1

    dev = torch.device("cuda:0")
    csr = fem.CSRStructure(od.element_dofs(), n, dev)
    g_rp, g_col = csr.to_host()
    assert np.array_equal(g_rp, rp) and np.array_equal(g_col, col)
    red = fem.CoupledReduction(csr, owner, coef, constrained)
    assert red.n_free == n_s and red.nnz == len(scol)
    assert np.array_equal(red.row_ptr.cpu().numpy(), srp)
    assert np.array_equal(red.col[: red.nnz].cpu().numpy(), scol)
    # row groups: the equations rows of each solver row, ascending
    gp = red.grp_ptr.cpu().numpy()
    gr = red.grp_rows.cpu().numpy()
    for s in range(n_s):
        assert np.array_equal(gr[gp[s]:gp[s + 1]], np.nonzero(owner == s)[0])
    v64 = torch.as_tensor(vals, device=dev)
    s64 = red.gather(v64, fem.F64).cpu().numpy()[: red.nnz]
    assert np.max(np.abs(s64 - sv)) <= 1e-14 * np.abs(sv).max()
    v32 = v64.float()
    s32 = red.gather(v32, fem.F32).cpu().numpy()[: red.nnz]
    _, _, sv32, _ = fo.reduce_coupled(rp, col, vals.astype(np.float32).astype(np.float64), owner, coef, ufix, src)
    assert np.max(np.abs(s32 - sv32)) <= 1e-6 * np.abs(sv32).max()
    u = torch.as_tensor(np.where(fixed, ufix, 0.0), device=dev)
    bg = red.rhs(v64, u, src=torch.as_tensor(src, device=dev)).cpu().numpy()[:n_s]
    assert np.max(np.abs(bg - b)) <= 1e-13 * np.abs(b).max()
    # solve on the SELL solver matrix (values gathered from the coupled sums)
    S = fem.SolverMatrix(red)
    ws = fem.KrylovWorkspace(n_s, dev)
    x = torch.zeros(n_s, dtype=torch.float64, device=dev)
    bt = torch.as_tensor(b, device=dev)
    info = fem.cg(S, S.gather(v64, fem.F64), bt, x, 1e-12, 10000, ws)
    red.field_update(x, u)
    torch.cuda.synchronize()
    A = _dense(srp, scol, sv, n_s)
    xn = x.cpu().numpy()
    assert np.linalg.norm(b - A @ xn) <= 1e-11 * np.linalg.norm(b), info
    u_ref = fo.field_update_coupled(owner, coef, np.linalg.solve(A, b), ufix)
    un = u.cpu().numpy()
    assert np.max(np.abs(un - u_ref)) <= 1e-9 * np.abs(u_ref).max()
    assert np.all(un[fixed] == ufix[fixed])
    for c, m, a in ties:
        assert un[c] == un[m] * a


def test_periodic_laplace_through_the_api_keeps_u_equal_x():
    from iron_amd import iron
    ne = (5, 4, 3)
    print(f"periodic Laplace API ne={ne}", flush=True)
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
                                        "linearIterativeAbsoluteTolerance": 0.0})
    rm = geo._mesh()
    X = rm.coords
    users = rm.comp_user[0]
    for i in np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]:
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), 1, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in np.nonzero(np.abs(X[:, 0] - 1) < 1e-12)[0]:
        bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), 1, iron.BoundaryConditionsTypes.FIXED, 1.0)
    for c, m, _ in _face_ties(X, lambda i: 1.0):
        bc.ConstrainNodeDofsEqual(dep, iron.FieldVariableTypes.U, 1, 1, 1, [int(users[m]), int(users[c])], 1.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    n_ties = len(_face_ties(X, lambda i: 1.0))
    assert es._device_cache.red.n_free == len(X) - 2 * (ne[1] + 1) * (ne[2] + 1) - n_ties
    assert np.max(np.abs(u - X[:, 0])) < 1e-9


def test_elasticity_end_plate_tie_mixed_precision_matches_the_oracle():
    """Cantilever clamped at x=0 with a unit z traction on x=3; the z-displacements of
    the end face are tied to one DOF (a plate that translates in z).  Mixed-precision
    CG (FP32 inner, FP64 refinement) on node-interleaved SELL rows."""
    from iron_amd import iron
    ne, extent = (6, 2, 2), (3.0, 1.0, 1.0)
    print(f"elasticity end-plate tie ne={ne}", flush=True)
    rng = np.random.default_rng(5)
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
        off = mat._offset(iron.FieldVariableTypes.U, c)
        mat.ParameterSetDataGet(iron.FieldVariableTypes.U, 1)[off:off + len(E)] = E
    _equations(iron, es)
    problem, solver, se, bc = _problem(iron, es, [iron.ProblemClasses.ELASTICITY,
                                                  iron.ProblemTypes.LINEAR_ELASTICITY,
                                                  iron.ProblemSubtypes.NONE],
                                       {"linearIterativeType": iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT,
                                        "linearIterativeRelativeTolerance": 1.0e-10,
                                        "linearIterativeAbsoluteTolerance": 0.0,
                                        "linearIterativePrecision": "mixed"})
    rm = geo._mesh()
    X = rm.coords
    users = rm.comp_user[0]
    clamp = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
    face = np.nonzero(np.abs(X[:, 0] - 3.0) < 1e-12)[0]
    for i in clamp:
        for c in (1, 2, 3):
            bc.SetNode(dep, iron.FieldVariableTypes.U, 1, 1, int(users[i]), c, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in face:
        bc.SetNode(dep, iron.FieldVariableTypes.DELUDELN, 1, 1, int(users[i]), 3,
                   iron.BoundaryConditionsTypes.NEUMANN_POINT, -1.0)
    bc.ConstrainNodeDofsEqual(dep, iron.FieldVariableTypes.U, 1, 1, 3, [int(users[i]) for i in face], 1.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u = dep.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    # oracle
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    nu = 0.3 * np.ones_like(E)
    C = fo.linear_elasticity_tensor(E, E, E, nu, nu, nu)
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs)
    n, N1 = od.num_dofs, len(X)
    fixed = np.zeros(n, bool)
    for c in range(3):
        fixed[c * N1 + clamp] = True
    pv = {int(2 * N1 + i): -1.0 for i in face}
    rhs_fixed = np.zeros(n, bool)
    rhs_fixed[list(pv)] = True
    f = fo.neumann_integrate(om, od, pv)
    zf = [int(2 * N1 + i) for i in face]
    owner, coef = fo.constraint_couplings(fixed, [(c, zf[0], 1.0) for c in zf[1:]])
    srp, scol, sv, b = fo.reduce_coupled(rp, col, vals, owner, coef, np.zeros(n), f, rhs_fixed)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(fo._csr_scipy(srp, scol, sv).tocsc(), b)
    u_ref = fo.field_update_coupled(owner, coef, x, np.zeros(n))
    assert np.max(np.abs(u - u_ref)) < 1e-8 * np.max(np.abs(u_ref))
    assert np.all(u[zf] == u[zf[0]]) and u[zf[0]] < 0
    assert solver.last_info.iterations > 0


def test_two_equations_sets_in_one_solver_system_match_separate_oracle_solves():
    """A Laplace set (with periodic ties) and a linear-elasticity set on two regions,
    added to one solver equations: one block-diagonal solver system and one CG solve;
    each set's field matches its own oracle solve."""
    from iron_amd import iron
    print("two equations sets in one solver system", flush=True)
    ne1 = (4, 3, 2)
    r1, d1, g1, m1 = _setup(iron, ne1, iron.BasisInterpolationSpecifications.QUADRATIC_LAGRANGE, (1.0, 1.0, 1.0))
    es1, dep1 = _equations_set(iron, r1, g1, [iron.EquationsSetClasses.CLASSICAL_FIELD,
                                              iron.EquationsSetTypes.LAPLACE_EQUATION,
                                              iron.EquationsSetSubtypes.STANDARD_LAPLACE])
    _equations(iron, es1)
    ne2, ext2 = (4, 2, 2), (2.0, 1.0, 1.0)
    rng = np.random.default_rng(4)
    E = 10.0 * np.exp(0.3 * rng.standard_normal(int(np.prod(ne2))))
    r2, d2, g2, m2 = _setup(iron, ne2, iron.BasisInterpolationSpecifications.LINEAR_LAGRANGE, ext2)
    es2, dep2 = _equations_set(iron, r2, g2, [iron.EquationsSetClasses.ELASTICITY,
                                              iron.EquationsSetTypes.LINEAR_ELASTICITY,
                                              iron.EquationsSetSubtypes.THREE_DIMENSIONAL])
    mat = iron.Field()
    es2.MaterialsCreateStart(4, mat)
    for c in (1, 2, 3):
        mat.ComponentInterpolationSet(iron.FieldVariableTypes.U, c, iron.FieldInterpolationTypes.ELEMENT_BASED)
    es2.MaterialsCreateFinish()
    for c in (1, 2, 3):
        off = mat._offset(iron.FieldVariableTypes.U, c)
        mat.ParameterSetDataGet(iron.FieldVariableTypes.U, 1)[off:off + len(E)] = E
    _equations(iron, es2)
    problem, solver, se, bc = _problem(iron, es1, [iron.ProblemClasses.CLASSICAL_FIELD,
                                                   iron.ProblemTypes.LAPLACE_EQUATION,
                                                   iron.ProblemSubtypes.STANDARD_LAPLACE],
                                       {"linearIterativeType": iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT,
                                        "linearIterativeRelativeTolerance": 1.0e-12,
                                        "linearIterativeAbsoluteTolerance": 0.0})
    assert se.EquationsSetAdd(es2) == 2
    X1, u1s = g1._mesh().coords, g1._mesh().comp_user[0]
    X2, u2s = g2._mesh().coords, g2._mesh().comp_user[0]
    for i in np.nonzero(np.abs(X1[:, 0]) < 1e-12)[0]:
        bc.SetNode(dep1, iron.FieldVariableTypes.U, 1, 1, int(u1s[i]), 1, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in np.nonzero(np.abs(X1[:, 0] - 1) < 1e-12)[0]:
        bc.SetNode(dep1, iron.FieldVariableTypes.U, 1, 1, int(u1s[i]), 1, iron.BoundaryConditionsTypes.FIXED, 1.0)
    ties = _face_ties(X1, lambda i: 1.0)
    for c, m, _ in ties:
        bc.ConstrainNodeDofsEqual(dep1, iron.FieldVariableTypes.U, 1, 1, 1, [int(u1s[m]), int(u1s[c])], 1.0)
    clamp = np.nonzero(np.abs(X2[:, 0]) < 1e-12)[0]
    face = np.nonzero(np.abs(X2[:, 0] - ext2[0]) < 1e-12)[0]
    for i in clamp:
        for c in (1, 2, 3):
            bc.SetNode(dep2, iron.FieldVariableTypes.U, 1, 1, int(u2s[i]), c, iron.BoundaryConditionsTypes.FIXED, 0.0)
    for i in face:
        bc.SetNode(dep2, iron.FieldVariableTypes.DELUDELN, 1, 1, int(u2s[i]), 3,
                   iron.BoundaryConditionsTypes.NEUMANN_POINT, -1.0)
    se.BoundaryConditionsCreateFinish()
    problem.Solve()
    u1 = dep1.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    u2 = dep2.ParameterSetDataGet(iron.FieldVariableTypes.U, 1).copy()
    blk = es1._block_cache.red
    assert blk.n_free == es1._device_cache.red.n_free + es2._device_cache.red.n_free
    # set 1: periodic Laplace keeps u = x (Q2 reproduces linear fields exactly)
    assert np.max(np.abs(u1 - X1[:, 0])) < 1e-9
    # set 2: its own oracle solve
    ob = fo.Basis(3, 1, 2)
    om = fo.regular_mesh(ne2, [ob], ext2)
    od = fo.DofMap(om, [0, 0, 0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    nu = 0.3 * np.ones_like(E)
    C = fo.linear_elasticity_tensor(E, E, E, nu, nu, nu)
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs)
    n, N1 = od.num_dofs, len(X2)
    fixed = np.zeros(n, bool)
    for c in range(3):
        fixed[c * N1 + clamp] = True
    pv = {int(2 * N1 + i): -1.0 for i in face}
    rhs_fixed = np.zeros(n, bool)
    rhs_fixed[list(pv)] = True
    f = fo.neumann_integrate(om, od, pv)
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, np.zeros(n), f, rhs_fixed)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(fo._csr_scipy(srp, scol, sv).tocsc(), b)
    assert np.max(np.abs(u2[free] - x)) < 1e-8 * np.max(np.abs(x))
    assert np.all(u2[fixed] == 0.0)
