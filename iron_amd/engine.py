"""Problem.Solve() for the iron-compatible API: maps an equations set, its
fields, boundary conditions and solver settings onto the device engine.

Static linear solve, following PROBLEM_SOLVER_EQUATIONS_SOLVE ->
Problem_SolverEquationsStaticLinearSolve (problem_routines.F90:2533):
  assemble (EQUATIONS_SET_ASSEMBLE) -> solver matrix + RHS
  (SOLVER_MATRICES_STATIC_ASSEMBLE) -> KSP -> field update -> back-substitution
  (EQUATIONS_SET_BACKSUBSTITUTE, problem_routines.F90:2597).
Device structures (CSR pattern, scatter map, reduction, solver matrix) are
cached on the equations set and reused while the mesh and the set of fixed
DOFs stay the same.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np
import torch

from . import fem
from ._lib import CMFEError, NotConvergedError
from .basis import LagrangeBasis, face_quadrature
from .mesh import DofMap, RegularMesh, structured_colours


def _iron():
    from . import iron
    return iron


# ------------------------------------------------------------ Neumann
def boundary_faces(rm: RegularMesh):
    """Boundary faces of a regular mesh: (element, xi direction, side 0/1)."""
    eidx = np.array(np.unravel_index(np.arange(rm.num_elements), rm.n_elem, order="F")).T
    out = []
    for k in range(rm.dim):
        for side in (0, 1):
            sel = eidx[:, k] == (0 if side == 0 else rm.n_elem[k] - 1)
            for e in np.nonzero(sel)[0]:
                out.append((int(e), k, side))
    return out


def neumann_matrix(rm: RegularMesh, dm: DofMap, comp_basis, point_dofs, geom_comp=0, integrated_only=frozenset(),
                   coords=None):
    """Neumann integration matrix N of BoundaryConditions_NeumannMatricesInitialise
    (boundary_condition_routines.F90:2227-2583): rows are the RHS variable's DOFs,
    columns the NEUMANN_POINT DOFs in the order of ``point_dofs``, and
    N[m, o] = sum over the boundary faces holding both nodes of int_face phi_m phi_o dA
    (3-D node-based case of BoundaryConditions_NeumannIntegrate, :2654-2996, for the
    face nodes m of o's component).  A face holding a NEUMANN_INTEGRATED_ONLY DOF of
    that component is skipped (:2860-2862).  Geometry is interpolated with the basis of
    mesh component ``geom_comp`` from the node coordinates ``coords`` ([nodes of that
    component, dim]; the geometric field's values, default the generated mesh's).
    Returns CSR (row_ptr [n+1], col, vals) with sorted, unique columns per row; f = N q
    is the integrated set (FIELD_INTEGRATED_NEUMANN_SET)."""
    n = dm.num_dofs
    point_dofs = np.asarray(point_dofs, dtype=np.int64)
    col_of = {int(d): j for j, d in enumerate(point_dofs)}
    rows, cols, vals = [], [], []
    coords = rm.coords if coords is None else np.asarray(coords, dtype=np.float64)
    comps = sorted(set(int(c) for c in dm.component_of(point_dofs))) if len(point_dofs) else []
    for c in comps:
        mc = dm.comp_mesh[c]
        b: LagrangeBasis = comp_basis[c]
        en = rm.elem_nodes[mc]
        geo_en = rm.elem_nodes[geom_comp]
        gb = rm.bases[geom_comp]
        last = b.nodes_xi - 1
        for (e, k, side) in boundary_faces(rm):
            fnodes = np.nonzero(b.node_index[:, k] == (0 if side == 0 else last))[0]
            dofs = dm.offsets[c] + en[e, fnodes]
            has = np.array([int(d) in col_of for d in dofs])
            if not has.any():
                continue
            if integrated_only and any(int(d) in integrated_only for d in dofs):
                continue
            N, _, wf = face_quadrature(b, k, side)
            _, dNg, _ = face_quadrature(gb, k, side)
            xe = coords[geo_en[e]]
            tang = [j for j in range(rm.dim) if j != k]
            dx = np.einsum("gaj,ad->gdj", dNg[:, :, tang], xe)  # [ng, dim, 2]
            gl = np.einsum("gdi,gdj->gij", dx, dx)
            jac = np.sqrt(np.abs(gl[:, 0, 0] * gl[:, 1, 1] - gl[:, 0, 1] * gl[:, 1, 0]))
            Mf = np.einsum("ga,gb,g->ab", N[:, fnodes], N[:, fnodes], wf * jac)
            src = np.nonzero(has)[0]
            rr, cc = np.meshgrid(dofs, [col_of[int(dofs[j])] for j in src], indexing="ij")
            rows.append(rr.ravel())
            cols.append(cc.ravel())
            vals.append(Mf[:, src].ravel())
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    if not rows:
        return row_ptr, np.zeros(0, dtype=np.int64), np.zeros(0)
    r = np.concatenate(rows).astype(np.int64)
    cidx = np.concatenate(cols).astype(np.int64)
    v = np.concatenate(vals)
    key = r * max(len(point_dofs), 1) + cidx
    order = np.argsort(key, kind="stable")
    uk, first = np.unique(key[order], return_index=True)
    sv = np.add.reduceat(v[order], first)
    ur = uk // max(len(point_dofs), 1)
    np.add.at(row_ptr, ur + 1, 1)
    return np.cumsum(row_ptr), uk % max(len(point_dofs), 1), sv


def neumann_integrate(rm: RegularMesh, dm: DofMap, comp_basis, point_values: dict, geom_comp=0,
                      integrated_only=frozenset(), coords=None):
    """Integrated Neumann values f = N q on the host (see neumann_matrix);
    ``point_values``: {dof: q}; ``integrated_only``: set of DOFs."""
    if not point_values:
        return np.zeros(dm.num_dofs)
    pd = np.array(sorted(point_values), dtype=np.int64)
    rp, col, v = neumann_matrix(rm, dm, comp_basis, pd, geom_comp, integrated_only, coords)
    q = np.array([point_values[int(d)] for d in pd])
    f = np.zeros(dm.num_dofs)
    rows = np.repeat(np.arange(dm.num_dofs), np.diff(rp))
    np.add.at(f, rows, v * q[col])
    return f


class NeumannIntegrator:
    """The integration matrix on the device: NeumannIntegrate's f = N q
    (boundary_condition_routines.F90:2654-2996) as one CSR product per solve, with
    the point values q resident in HBM."""

    def __init__(self, rm: RegularMesh, dm: DofMap, comp_basis, point_dofs, device, geom_comp=0,
                 integrated_only=frozenset(), coords=None):
        from . import _lib
        self.point_dofs = np.asarray(point_dofs, dtype=np.int64)
        rp, col, v = neumann_matrix(rm, dm, comp_basis, self.point_dofs, geom_comp, integrated_only, coords)
        self.n = dm.num_dofs
        self.nnz = len(col)
        self.row_ptr = torch.as_tensor(rp.astype(np.int32), device=device)
        self.col = torch.as_tensor(np.maximum(col, 0).astype(np.int32) if len(col) else np.zeros(1, np.int32),
                                   device=device)
        self.vals = torch.as_tensor(v if len(v) else np.zeros(1), dtype=torch.float64, device=device)
        self.desc = _lib.MatrixDesc(self.n, _lib.MIFE_FMT_CSR, _lib.ptr(self.row_ptr), _lib.ptr(self.col), None, 0,
                                    self.nnz, None)

    def integrate(self, q: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """f [n] = N q for point values q [len(point_dofs)] (FP64, device)."""
        from . import _lib
        if out is None:
            out = torch.empty(self.n, dtype=torch.float64, device=q.device)
        if len(self.point_dofs) == 0:
            return out.zero_()
        _lib.call("mife_spmv", C.byref(self.desc), _lib.ptr(self.vals), _lib.ptr(q), _lib.ptr(out), _lib.MIFE_F64,
                  _lib.stream_ptr())
        return out


# ------------------------------------------------------------ helpers
def geometry_coords(geo, rm: RegularMesh, mesh_comp: int) -> np.ndarray:
    """Node coordinates [nodes, dim] of the geometric field, whose components must be
    node based on mesh component ``mesh_comp`` (0-based): the values the element
    kernels interpolate geometry from (iron interpolates the geometric field,
    FIELD_INTERPOLATE_XI on the GEOMETRIC_FIELD's parameters, field_routines.F90:8042-8260)."""
    iron = _iron()
    var = geo._var(iron.FieldVariableTypes.U)
    if var.ncomp < rm.dim:
        raise CMFEError(1, f"the geometric field needs {rm.dim} components")
    for k in range(rm.dim):
        if var.mesh_comp[k] - 1 != mesh_comp or var.interp[k] != iron.FieldInterpolationTypes.NODE_BASED:
            raise CMFEError(1, f"geometric field component {k + 1} must be node based on mesh component "
                               f"{mesh_comp + 1}")
    return np.ascontiguousarray(np.stack([geo._component(iron.FieldVariableTypes.U, k + 1) for k in range(rm.dim)],
                                         axis=1), dtype=np.float64)


def _dof_map(field, rm: RegularMesh):
    iron = _iron()
    v = field._var(iron.FieldVariableTypes.U)
    return DofMap(rm, [mc - 1 for mc in v.mesh_comp])


def _element_values(field, comp: int, rm: RegularMesh, vt=None):
    """Per-element value of a field component (element-based or constant)."""
    iron = _iron()
    vt = vt or iron.FieldVariableTypes.U
    var = field._var(vt)
    it = var.interp[comp - 1]
    vals = field._component(vt, comp)
    if it == iron.FieldInterpolationTypes.ELEMENT_BASED:
        return np.ascontiguousarray(vals, dtype=np.float64)
    if it == iron.FieldInterpolationTypes.CONSTANT:
        return np.full(rm.num_elements, float(vals[0]))
    raise CMFEError(1, "node-based material parameters are not supported by the element kernels")


def _is(a, b):
    return a is b or a.__dict__ is b.__dict__


def _boundary_conditions(bc, es, dm: DofMap, rm: RegularMesh, others=()):
    """Conditions on this set's dependent field; entries on the dependent fields in
    ``others`` (the other sets of the same solver equations) are skipped."""
    iron = _iron()
    dep = es.dependentField
    n = dm.num_dofs
    fixed = np.zeros(n, dtype=bool)
    rhs_fixed = np.zeros(n, dtype=bool)
    u = dep._data(iron.FieldVariableTypes.U).copy()
    neumann = {}
    increments = {}
    integrated_only = set()
    for (field, vt, node, comp, cond, value, add) in bc._entries:
        if any(_is(field, o) for o in others):
            continue
        if not _is(field, dep):
            raise CMFEError(1, "boundary conditions must be set on the dependent field")
        mc = dm.comp_mesh[comp - 1]
        dof = int(dm.offsets[comp - 1] + rm.local_node(node, mc))
        if vt == iron.FieldVariableTypes.U:
            if cond in (iron.BoundaryConditionsTypes.FIXED, iron.BoundaryConditionsTypes.FIXED_INCREMENTED):
                fixed[dof] = True
                u[dof] = (u[dof] + value) if add else value
                if cond == iron.BoundaryConditionsTypes.FIXED_INCREMENTED:
                    increments[dof] = value
            else:
                raise CMFEError(1, f"boundary condition type {cond} on U is not supported")
        elif vt == iron.FieldVariableTypes.DELUDELN:
            # both make the RHS DOF "fixed" (boundary_condition_routines.F90:1685-1693)
            rhs_fixed[dof] = True
            if cond == iron.BoundaryConditionsTypes.NEUMANN_POINT:
                neumann[dof] = (neumann.get(dof, 0.0) + value) if add else value
            elif cond == iron.BoundaryConditionsTypes.FIXED:
                dep._data(iron.FieldVariableTypes.DELUDELN)[dof] = value
            elif cond in (iron.BoundaryConditionsTypes.NEUMANN_INTEGRATED,
                          iron.BoundaryConditionsTypes.NEUMANN_INTEGRATED_ONLY):
                # the integrated value goes straight into the RHS values (:1366-1370)
                dep._data(iron.FieldVariableTypes.DELUDELN)[dof] += value
                if cond == iron.BoundaryConditionsTypes.NEUMANN_INTEGRATED_ONLY:
                    integrated_only.add(dof)
            else:
                raise CMFEError(1, f"boundary condition type {cond} on DELUDELN is not supported")
    return fixed, u, neumann, increments, rhs_fixed, frozenset(integrated_only)


def _dof_constraints(bc, es, dm: DofMap, rm: RegularMesh, fixed: np.ndarray, others=()):
    """Solver DOF and coupling coefficient of every DOF under the constraints of
    ConstrainNodeDofsEqual (DofConstraintSet / DofConstraintsCreateFinish,
    boundary_condition_routines.F90:3204-3472; solver rows of SOLVER_MAPPING_CALCULATE,
    solver_mapping_routines.F90:422-520).  Returns (owner, coef, constrained) or None."""
    iron = _iron()
    cons = [k for k in getattr(bc, "_constraints", []) if not any(_is(k[0], o) for o in others)]
    if not cons:
        return None
    dep = es.dependentField
    n = dm.num_dofs
    master = np.full(n, -1, dtype=np.int64)
    coef = np.ones(n, dtype=np.float64)
    for (field, vt, comp, nodes, a) in cons:
        if any(_is(field, o) for o in others):
            continue
        if not _is(field, dep):
            raise CMFEError(1, "DOF constraints must be set on the dependent field")
        if vt != iron.FieldVariableTypes.U:
            raise CMFEError(1, "DOF constraints are implemented on the U variable")
        mc = dm.comp_mesh[comp - 1]
        dofs = [int(dm.offsets[comp - 1] + rm.local_node(nd, mc)) for nd in nodes]
        m = dofs[0]
        for c in dofs[1:]:
            if master[c] >= 0:
                raise CMFEError(1, f"DOF {c + 1} is constrained twice")
            master[c] = m
            coef[c] = a
    constrained = master >= 0
    if np.any(fixed & constrained):
        raise CMFEError(1, "a DOF that is part of a linear constraint also has a fixed boundary condition")
    if np.any(fixed[master[constrained]]) or np.any(constrained[master[constrained]]):
        raise CMFEError(1, "a DOF depended on by a linear constraint is not free in the boundary conditions")
    solver_dof = ~fixed & ~constrained
    rank = np.full(n, -1, dtype=np.int64)
    rank[solver_dof] = np.arange(int(solver_dof.sum()))
    owner = np.where(constrained, rank[np.maximum(master, 0)], rank)
    return owner.astype(np.int32), coef, constrained


class _LinearCache:
    pass


def _linear_setup(es, rm: RegularMesh, dm: DofMap, fixed: np.ndarray, couple=None):
    """Device structures for a linear equations set, cached while valid."""
    cache = getattr(es, "_device_cache", None)
    key = (id(rm), dm.num_dofs, fixed.tobytes(),
           None if couple is None else (couple[0].tobytes(), couple[1].tobytes(), couple[2].tobytes()))
    if cache is not None and cache.key == key:
        return cache
    c = _LinearCache()
    c.key = key
    dev = torch.device("cuda:0")
    ed = dm.element_dofs()
    c.csr = fem.CSRStructure(ed, dm.num_dofs, dev, colours=structured_colours(rm))
    if couple is None:
        c.red = fem.DirichletReduction(c.csr, fixed)
        c.lead = np.nonzero(~fixed)[0]
    else:
        c.red = fem.CoupledReduction(c.csr, *couple)
        c.lead = c.red.lead
    # a vector field on one mesh component: rows node by node (16-bit SELL columns);
    # three components whose free DOFs are whole nodes: 3x3 node blocks (SELL3)
    ncomp = len(dm.comp_mesh)
    nper = rm.num_nodes(dm.comp_mesh[0])
    inter = (ncomp, nper) if ncomp > 1 and dm.num_dofs == ncomp * nper else None
    c.S = fem.SolverMatrix(c.red, interleave=inter, block=ncomp == 3)
    c.ws = fem.KrylovWorkspace(c.red.n_free, dev)
    es._device_cache = c
    return c


# ---------------------------------------------------------------- solve
def solve(problem, es, solver, bc):
    kind = es._kind
    if kind == "mooney_rivlin":
        from . import finite_elasticity
        return finite_elasticity.solve(problem, es, solver, bc)
    return solve_linear(es, solver, bc)


def _prepare_linear(es, solver, bc, others=(), force_prec=None):
    """Structures, assembly and the RHS pieces of one linear equations set
    (``force_prec`` overrides the solver's precision setting)."""
    iron = _iron()
    geo = es.geometricField
    rm: RegularMesh = geo._mesh()
    dep = es.dependentField
    dm = _dof_map(dep, rm)
    mcs = set(dm.comp_mesh)
    if len(mcs) != 1:
        raise CMFEError(1, "linear equations sets need every dependent component on one mesh component")
    mc = dm.comp_mesh[0]
    b: LagrangeBasis = rm.bases[mc]
    if mc != 0:
        raise CMFEError(1, "the dependent field must use the geometric mesh component")
    fixed, u_host, neumann, _, rhs_fixed, integrated_only = _boundary_conditions(bc, es, dm, rm, others)
    c = _linear_setup(es, rm, dm, fixed, _dof_constraints(bc, es, dm, rm, fixed, others))
    dev = c.csr.device
    coords = geometry_coords(geo, rm, 0)
    kind = es._kind
    prec = force_prec or str(getattr(solver, "linearIterativePrecision", "fp64")).lower()
    if prec not in ("fp64", "mixed", "auto"):
        raise CMFEError(1, f"linearIterativePrecision must be 'fp64', 'mixed' or 'auto', not {prec!r}")
    if prec == "auto":
        prec = fem.auto_precision(len(dm.comp_mesh), c.S.stored)
    mixed = prec == "mixed"
    solver.linearIterativePrecisionUsed = prec
    # the mixed solve's FP32 values are the FP64 values rounded, taken in the
    # same gather as the FP64 solver values: one assembly, in FP64
    v64 = c.csr.zeros(fem.F64)
    src = None
    # generated regular meshes in 3-D take the structured assembly (no position map)
    grid = rm.n_elem if rm.dim == 3 else None
    if kind in ("laplace", "generalised_laplace", "poisson"):
        asm = fem.ScalarAssembler(b, coords, rm.elem_nodes[mc], c.csr, grid=grid)
        kappa = None
        source = None
        if kind == "generalised_laplace":
            mf = es.materialsField
            k11, k22, k33 = (_element_values(mf, j, rm) for j in (1, 2, 3))
            offd = [_element_values(mf, j, rm) for j in (4, 5, 6)] if mf._var(iron.FieldVariableTypes.U).ncomp >= 6 \
                else []
            if not (np.allclose(k11, k22) and np.allclose(k11, k33) and all(np.all(o == 0) for o in offd)):
                raise CMFEError(1, "only isotropic conductivity (k11 = k22 = k33, no off-diagonals) is supported")
            kappa = torch.as_tensor(k11, device=dev)
            if es.sourceField is not None:
                source = torch.as_tensor(_element_values(es.sourceField, 1, rm), device=dev)
        elif kind == "poisson":
            source = torch.as_tensor(_element_values(es.materialsField, 1, rm), device=dev)
        if source is not None:
            src = torch.zeros(dm.num_dofs, dtype=torch.float64, device=dev)
        asm.assemble(v64, kappa=kappa, source=source, src_vec=src, prec=fem.F64)
    elif kind == "linear_elasticity":
        if rm.dim != 3:
            raise CMFEError(1, "linear elasticity is implemented for the 3-D subtype")
        asm = fem.ElasticAssembler(b, coords, rm.elem_nodes[mc], c.csr, grid=grid)
        mf = es.materialsField
        mat = np.stack([_element_values(mf, j, rm) for j in range(1, 7)], axis=1)
        mt = torch.as_tensor(np.ascontiguousarray(mat), device=dev)
        asm.assemble(v64, mt, prec=fem.F64)
    else:
        raise CMFEError(1, f"equations set kind {kind} is not implemented")
    # RHS: DELUDELN values of rows whose RHS DOF is fixed, plus the integrated
    # Neumann point values (b = f + N q, solver_routines.F90:14681-14696,14850)
    rhs = torch.as_tensor(np.where(rhs_fixed, dep._data(iron.FieldVariableTypes.DELUDELN), 0.0), device=dev)
    if neumann:
        # integration matrix built once per set of point DOFs, f = N q on the device
        pd = np.array(sorted(neumann), dtype=np.int64)
        nkey = (pd.tobytes(), tuple(sorted(integrated_only)), coords.tobytes())
        if getattr(c, "neumann_key", None) != nkey:
            c.neumann = NeumannIntegrator(rm, dm, [b] * len(dm.comp_mesh), pd, dev, integrated_only=integrated_only,
                                          coords=coords)
            c.neumann_key = nkey
        q = torch.as_tensor(np.array([neumann[int(d)] for d in pd], dtype=np.float64), device=dev)
        rhs += c.neumann.integrate(q)
    rfix = torch.as_tensor(rhs_fixed.astype(np.uint8), device=dev) if rhs_fixed.any() else None
    u = torch.as_tensor(np.where(fixed, u_host, 0.0), device=dev)
    x0 = u_host[c.lead] if c.red.n_free else np.zeros(0)
    return dict(es=es, dep=dep, c=c, v64=v64, src=src, rhs=rhs, rfix=rfix, u=u, x0=x0, mixed=mixed)


def _krylov(S, red, vals64, bvec, x, ws, solver, mixed):
    """KSP settings (solver_routines.F90:10921-11157) and the solve."""
    iron = _iron()
    rtol = float(solver.linearIterativeRelativeTolerance)
    atol = float(solver.linearIterativeAbsoluteTolerance)
    maxit = int(solver.linearIterativeMaximumIterations)
    if int(solver.linearType) == iron.LinearSolverTypes.DIRECT:
        rtol, atol = 1e-14, 0.0  # no GPU direct factorisation: iterate to round-off
    bn = float(torch.linalg.vector_norm(bvec).item()) if red.n_free else 0.0
    rt = max(rtol, atol / bn) if bn > 0 else rtol
    ktype = int(solver.linearIterativeType)
    if red.n_free == 0:
        return fem.SolveInfo(0, 0.0)
    try:
        return _krylov_run(S, vals64, bvec, x, ws, solver, mixed, ktype, rt, maxit)
    except NotConvergedError as e:
        # iron flags a warning and keeps the iterate (solver_routines.F90:11875-11880);
        # the library's message names the actual stop (iteration limit, or a true
        # residual above the tolerance after the recurrence's restarts)
        info = getattr(e, "info", None) or fem.SolveInfo(maxit, float("nan"))
        what = "direct (iterated to round-off)" if int(solver.linearType) == iron.LinearSolverTypes.DIRECT \
            else "iterative"
        warnings.warn(f"Linear {what} solver did not converge: {e}", RuntimeWarning, stacklevel=3)
        return info


def _krylov_run(S, vals64, bvec, x, ws, solver, mixed, ktype, rt, maxit):
    iron = _iron()
    if mixed and ktype == iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT:
        s64, s32 = S.gather2(vals64)
        return fem.refine(S, s64, s32, bvec, x, rt, maxit, ws)
    s64 = S.gather(vals64, fem.F64)
    if ktype == iron.IterativeLinearSolverTypes.CONJUGATE_GRADIENT:
        return fem.cg(S, s64, bvec, x, rt, maxit, ws)
    if ktype == iron.IterativeLinearSolverTypes.GMRES or int(solver.linearType) == iron.LinearSolverTypes.DIRECT:
        return fem.gmres(S, s64, bvec, x, rt, maxit, ws, restart=int(solver.linearIterativeGMRESRestart))
    raise CMFEError(1, f"iterative solver type {ktype} is not implemented (CG and GMRES are)")


def _finish_linear(st, info):
    """Back-substitution (iron's DELUDELN values and the correct reactions) and the
    field writes of one set after its field update."""
    iron = _iron()
    c, v64, u, src = st["c"], st["v64"], st["u"], st["src"]
    R_iron = fem.backsubstitute(c.csr, v64, u, src=src, rhs_fixed=st["rfix"], iron_compat=True)
    R = fem.backsubstitute(c.csr, v64, u, src=src, iron_compat=False)
    dep, es = st["dep"], st["es"]
    dep._data(iron.FieldVariableTypes.U)[:] = u.cpu().numpy()
    dep._data(iron.FieldVariableTypes.DELUDELN)[:] = R_iron.cpu().numpy()
    es.reactions = R.cpu().numpy()
    es._last = dict(vals=v64, src=src, u=u, info=info)


def solve_linear(es, solver, bc):
    st = _prepare_linear(es, solver, bc)
    c = st["c"]
    bvec = c.red.rhs(st["v64"], st["u"], src=st["src"], rhs=st["rhs"], rhs_fixed=st["rfix"])
    x = torch.as_tensor(st["x0"], device=c.csr.device).clone()
    info = _krylov(c.S, c.red, st["v64"], bvec, x, c.ws, solver, st["mixed"])
    solver.last_info = info
    c.red.field_update(x, st["u"])
    _finish_linear(st, info)
    return info


class _BlockCache:
    pass


def solve_linear_sets(sets, solver, bc):
    """Several linear equations sets in one solver equations: one block-diagonal
    solver system (rows and columns set by set, solver_mapping_routines.F90:791-830),
    one Krylov solve, then each set's field update and back-substitution."""
    sts = []
    # one precision for the joint system: "auto" keeps several sets in FP64
    force = "fp64" if str(getattr(solver, "linearIterativePrecision", "fp64")).lower() == "auto" else None
    for es in sets:
        others = tuple(o.dependentField for o in sets if o is not es)
        sts.append(_prepare_linear(es, solver, bc, others, force_prec=force))
    mixed = sts[0]["mixed"]
    key = tuple(id(st["c"].red) for st in sts)  # the block cache holds these reductions
    holder = sets[0]
    bc_cache = getattr(holder, "_block_cache", None)
    if bc_cache is None or bc_cache.key != key:
        bc_cache = _BlockCache()
        bc_cache.key = key
        bc_cache.red = fem.BlockReduction([st["c"].red for st in sts])
        bc_cache.S = fem.SolverMatrix(bc_cache.red)
        bc_cache.ws = fem.KrylovWorkspace(bc_cache.red.n_free, sts[0]["c"].csr.device)
        holder._block_cache = bc_cache
    red, S = bc_cache.red, bc_cache.S
    bvec = red.rhs([(st["v64"], st["u"], st["src"], st["rhs"], st["rfix"]) for st in sts])
    x = torch.as_tensor(np.concatenate([st["x0"] for st in sts]), device=sts[0]["c"].csr.device).clone()
    info = _krylov(S, red, [st["v64"] for st in sts], bvec, x, bc_cache.ws, solver, mixed)
    solver.last_info = info
    red.field_update(x, [st["u"] for st in sts])
    for st in sts:
        _finish_linear(st, info)
    return info
