"""OpenCMISS-Iron-compatible Python API over the MI355X engine.

Mirrors the classes, methods, properties and enum values of iron's generated
Python bindings (bindings/generate_bindings/python.py; enum values from
src/opencmiss_iron.F90 and the *_constants / *_routines modules they alias) for
the subset the static linear solve and the finite-elasticity load-increment
solve use, so that scripts such as bindings/python/tests/Laplace.py run
unchanged with ``from iron_amd import iron``.  ``problem.Solve()`` assembles
and solves on the GPU through libmi_fe.so.

Every ``XxxSet(value)`` / ``XxxGet()`` method is also available as the
property ``xxx`` (python.py:142-169 generates both forms).  Errors raise
CMFEError, as the bindings' wrapper does (_utils.in.py:66-77).

Extensions beyond iron (no counterpart in the reference):
  * ``Solver.linearIterativePrecision`` = "fp64" (default), "mixed" (FP32 inner
    CG with FP64 residual replacement) or "auto" (mixed for a scalar field whose
    solver matrix holds at least fem.MIXED_MIN_STORED entries, else FP64; the
    choice is left in ``Solver.linearIterativePrecisionUsed``);
  * a source field on a GENERALISED_LAPLACE equations set
    (``EquationsSet.SourceCreateStart``), which combines the element
    conductivity of the generalised Laplace subtype with a Poisson source
    vector, the diffusion problem the benchmark solves;
  * ``EquationsSet.reactions``: the correct K u - f on the RHS variable, next to
    the iron-compatible DELUDELN values written by Solve().
"""
from __future__ import annotations

import enum

import numpy as np

from ._lib import CMFEError  # noqa: F401  (re-exported as iron.CMFEError)


# ----------------------------------------------------------------- enums
class BasisTypes(enum.IntEnum):
    LAGRANGE_HERMITE_TP = 1
    SIMPLEX = 2


class BasisInterpolationSpecifications(enum.IntEnum):
    LINEAR_LAGRANGE = 1
    QUADRATIC_LAGRANGE = 2
    CUBIC_LAGRANGE = 3


class GeneratedMeshTypes(enum.IntEnum):
    REGULAR = 1


class GeneratedMeshSurfaceTypes(enum.IntEnum):
    # generated_mesh_routines.F90:103-108
    REGULAR_LEFT = 8
    REGULAR_RIGHT = 9
    REGULAR_TOP = 10
    REGULAR_BOTTOM = 11
    REGULAR_FRONT = 12
    REGULAR_BACK = 13


class DecompositionTypes(enum.IntEnum):
    ALL = 1
    CALCULATED = 2


class FieldTypes(enum.IntEnum):
    GEOMETRIC = 1
    FIBRE = 2
    GENERAL = 3
    MATERIAL = 4


class FieldVariableTypes(enum.IntEnum):
    U = 1
    DELUDELN = 2
    V = 5


class FieldParameterSetTypes(enum.IntEnum):
    VALUES = 1


class FieldDOFOrderTypes(enum.IntEnum):
    SEPARATED = 1
    CONTIGUOUS = 2


class FieldInterpolationTypes(enum.IntEnum):
    CONSTANT = 1
    ELEMENT_BASED = 2
    NODE_BASED = 3


class FieldScalingTypes(enum.IntEnum):
    NONE = 0
    UNIT = 1
    ARITHMETIC_MEAN = 3


class EquationsSetClasses(enum.IntEnum):
    ELASTICITY = 1
    CLASSICAL_FIELD = 4


class EquationsSetTypes(enum.IntEnum):
    # classical field class
    LAPLACE_EQUATION = 1
    POISSON_EQUATION = 2
    # elasticity class (same numbers, other class)
    LINEAR_ELASTICITY = 1
    FINITE_ELASTICITY = 2


class EquationsSetSubtypes(enum.IntEnum):
    NONE = 0
    STANDARD_LAPLACE = 1
    GENERALISED_LAPLACE = 2
    CONSTANT_SOURCE_POISSON = 1
    THREE_DIMENSIONAL = 4
    MOONEY_RIVLIN = 1
    NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN = 22


class ProblemClasses(enum.IntEnum):
    ELASTICITY = 1
    CLASSICAL_FIELD = 4


class ProblemTypes(enum.IntEnum):
    LAPLACE_EQUATION = 1
    POISSON_EQUATION = 2
    LINEAR_ELASTICITY = 1
    FINITE_ELASTICITY = 2


class ProblemSubtypes(enum.IntEnum):
    NONE = 0
    STANDARD_LAPLACE = 1
    GENERALISED_LAPLACE = 2
    LINEAR_SOURCE_POISSON = 1


class BoundaryConditionsTypes(enum.IntEnum):
    FREE = 0
    FIXED = 1
    NEUMANN_POINT = 8
    NEUMANN_INTEGRATED = 9
    FIXED_INCREMENTED = 13
    NEUMANN_INTEGRATED_ONLY = 20
    LINEAR_CONSTRAINT = 21


class LinearSolverTypes(enum.IntEnum):
    DIRECT = 1
    ITERATIVE = 2


class IterativeLinearSolverTypes(enum.IntEnum):
    RICHARDSON = 1
    CHEBYSHEV = 2
    CONJUGATE_GRADIENT = 3
    BICGSTAB = 4
    GMRES = 5


class IterativePreconditionerTypes(enum.IntEnum):
    NO = 0
    JACOBI = 1


class ControlLoopIdentifiers(enum.IntEnum):
    NODE = 0


class EquationsSparsityTypes(enum.IntEnum):
    SPARSE = 1
    FULL = 2


class SolverEquationsSparsityTypes(enum.IntEnum):
    SPARSE = 1
    FULL = 2


class EquationsOutputTypes(enum.IntEnum):
    NONE = 0
    TIMING = 1
    MATRIX = 2
    ELEMENT_MATRIX = 3


class SolverOutputTypes(enum.IntEnum):
    NONE = 0
    MONITOR = 1
    PROGRESS = 2
    TIMING = 3
    SOLVER = 4
    MATRIX = 5


class JacobianCalculationTypes(enum.IntEnum):
    FD = 1
    EQUATIONS = 2


class NewtonConvergenceTypes(enum.IntEnum):
    """SOLVER_NEWTON_CONVERGENCE_* (solver_routines.F90:250-252)."""
    PETSC_DEFAULT = 1
    ENERGY_NORM = 2
    DIFFERENTIATED_RATIO = 3


class NewtonLineSearchTypes(enum.IntEnum):
    """SOLVER_NEWTON_LINESEARCH_* (solver_routines.F90:231-234)."""
    NONORMS = 1
    LINEAR = 2
    QUADRATIC = 3
    CUBIC = 4


class SolverLibraries(enum.IntEnum):
    CMISS = 1
    PETSC = 2
    MUMPS = 3
    LAPACK = 11


class DiagnosticTypes(enum.IntEnum):
    ALL = 1
    IN = 2
    FROM = 3


# ---------------------------------------------------------- base object
class _Obj:
    """Created by CreateStart/CreateFinish; XxxSet / XxxGet map to properties."""

    def __getattr__(self, name):
        for suffix, kind in (("Set", "set"), ("Get", "get")):
            if name.endswith(suffix) and len(name) > 3 and name[0].isupper():
                prop = name[0].lower() + name[1:-3]
                if kind == "set":
                    return lambda value: setattr(self, prop, value)
                return lambda: getattr(self, prop)
        raise AttributeError(name)

    def Finalise(self):
        pass

    def Destroy(self):
        pass


_STATE = {"initialised": True}


def Initialise(*_args):
    _STATE["initialised"] = True
    return 0


def Finalise():
    _STATE["initialised"] = False


def ComputationalNumberOfNodesGet():
    return 1


def ComputationalNodeNumberGet():
    return 0


def DiagnosticsSetOn(*_args):
    pass


def ErrorHandlingModeSet(*_args):
    pass


def OutputSetOn(*_args):
    pass


# ------------------------------------------------- coordinates / regions
class CoordinateSystem(_Obj):
    def __init__(self):
        self.dimension = 3
        self.userNumber = None

    def CreateStart(self, userNumber):
        self.userNumber = userNumber

    def CreateFinish(self):
        pass


class Nodes(_Obj):
    def __init__(self):
        self.numberOfNodes = 0


class Region(_Obj):
    def __init__(self):
        self.label = ""
        self.coordinateSystem = None
        self.userNumber = None
        self._nodes = 0

    def CreateStart(self, userNumber, parent):
        self.userNumber = userNumber

    def CreateFinish(self):
        pass

    def NodesGet(self, nodes: Nodes):
        nodes.numberOfNodes = self._nodes


WorldCoordinateSystem = CoordinateSystem()
WorldRegion = Region()


# ---------------------------------------------------------------- basis
class Basis(_Obj):
    def __init__(self):
        self.type = BasisTypes.LAGRANGE_HERMITE_TP
        self.numberOfXi = 3
        self.interpolationXi = [BasisInterpolationSpecifications.LINEAR_LAGRANGE] * 3
        self.quadratureNumberOfGaussXi = None
        self._basis = None

    def CreateStart(self, userNumber):
        self.userNumber = userNumber

    def CreateFinish(self):
        from .basis import basis as make_basis
        if int(self.type) != BasisTypes.LAGRANGE_HERMITE_TP:
            raise CMFEError(1, "only Lagrange tensor-product bases are supported")
        interp = [int(i) for i in self.interpolationXi][: self.numberOfXi]
        if len(set(interp)) != 1:
            raise CMFEError(1, "mixed interpolation per xi direction is not supported")
        gx = None
        if self.quadratureNumberOfGaussXi is not None:
            g = [int(v) for v in self.quadratureNumberOfGaussXi][: self.numberOfXi]
            if len(set(g)) != 1:
                raise CMFEError(1, "anisotropic Gauss schemes are not supported")
            gx = g[0]
        self._basis = make_basis(self.numberOfXi, interp[0], gx)


# ------------------------------------------------------- mesh generation
class Mesh(_Obj):
    def __init__(self):
        self._regular = None


class GeneratedMesh(_Obj):
    def __init__(self):
        self.type = GeneratedMeshTypes.REGULAR
        self.basis = []
        self.extent = None
        self.numberOfElements = None
        self.origin = None
        self._region = None
        self._mesh = None

    def CreateStart(self, userNumber, region):
        self._region = region

    def CreateFinish(self, meshUserNumber, mesh: Mesh):
        from .mesh import RegularMesh
        bases = [b._basis for b in (self.basis if isinstance(self.basis, (list, tuple)) else [self.basis])]
        dim = bases[0].num_xi
        ne = [int(n) for n in self.numberOfElements][:dim]
        extent = [float(e) for e in (self.extent or [1.0] * dim)][:dim]
        origin = [float(o) for o in self.origin][:dim] if self.origin is not None else None
        mesh._regular = RegularMesh(ne, bases, extent, origin)
        mesh.numberOfComponents = len(bases)
        self._mesh = mesh
        if self._region is not None:
            self._region._nodes = mesh._regular.num_nodes(0)

    def SurfaceGet(self, *args):
        """SurfaceGet(surfaceType) or SurfaceGet(meshComponent, surfaceType) ->
        (surfaceNodes, normalXi): the user numbers of the mesh component's nodes on a
        face of the regular mesh and the signed xi direction of its outward normal
        (GENERATED_MESH_REGULAR_SURFACE_GET, generated_mesh_routines.F90:3943-4084).
        Nodes run with xi1 fastest, then xi2, then xi3 over the face, which on the
        generated lattice is ascending user number.  Directions beyond the mesh's
        dimension hold one layer of nodes, as in the reference."""
        if len(args) == 1:
            mc, st = 1, int(args[0])
        elif len(args) == 2:
            mc, st = int(args[0]), int(args[1])
        else:
            raise CMFEError(1, "SurfaceGet takes a surface type and optionally a mesh component first")
        rm = self._mesh._regular
        if not 1 <= mc <= len(rm.bases):
            raise CMFEError(1, f"mesh component {mc} does not exist")
        faces = {GeneratedMeshSurfaceTypes.REGULAR_LEFT: (0, False, -1),
                 GeneratedMeshSurfaceTypes.REGULAR_RIGHT: (0, True, 1),
                 GeneratedMeshSurfaceTypes.REGULAR_FRONT: (1, False, -2),
                 GeneratedMeshSurfaceTypes.REGULAR_BACK: (1, True, 2),
                 GeneratedMeshSurfaceTypes.REGULAR_BOTTOM: (2, False, -3),
                 GeneratedMeshSurfaceTypes.REGULAR_TOP: (2, True, 3)}
        if st not in faces:
            raise CMFEError(1, f"The specified surface type of {st} is invalid for a regular mesh.")
        axis, high, normal = faces[GeneratedMeshSurfaceTypes(st)]
        users = rm.comp_user[mc - 1]
        if axis >= rm.dim:
            return users.astype(np.int32).copy(), normal
        pos = rm.node_grid_pos[users - 1, axis]
        edge = rm.node_grid_pos[:, axis].max() if high else 0
        return users[pos == edge].astype(np.int32), normal

    def GeometricParametersCalculate(self, field):
        rm = self._mesh._regular
        data = field._data(FieldVariableTypes.U)
        n = rm.num_nodes(0)
        for c in range(rm.dim):
            data[field._offset(FieldVariableTypes.U, c + 1): field._offset(FieldVariableTypes.U, c + 1) + n] = \
                rm.coords[:, c]


class Decomposition(_Obj):
    def __init__(self):
        self.type = DecompositionTypes.CALCULATED
        self.numberOfDomains = 1
        self._mesh = None

    def CreateStart(self, userNumber, mesh):
        self._mesh = mesh

    def CreateFinish(self):
        pass

    def NodeDomainGet(self, nodeUserNumber, meshComponent):
        return 0

    def ElementDomainGet(self, elementUserNumber):
        return 0


# ---------------------------------------------------------------- fields
class _Variable:
    def __init__(self, ncomp):
        self.ncomp = ncomp
        self.mesh_comp = [1] * ncomp
        self.interp = [FieldInterpolationTypes.NODE_BASED] * ncomp
        self.label = ""
        self.values = None   # FP64 numpy, separated order
        self.offsets = None


class Field(_Obj):
    """A field with one or more variables; values live on the host in
    separated component order (one DOF per node for node-based components,
    one per element for element-based, one for constant)."""

    def __init__(self):
        self.type = FieldTypes.GEOMETRIC
        self.meshDecomposition = None
        self.geometricField = None
        self.fieldScalingType = FieldScalingTypes.NONE
        self.numberOfVariables = 1
        self._vars = {FieldVariableTypes.U: _Variable(3)}
        self._finished = False
        self._dof_order = {}

    # -- creation
    def CreateStart(self, userNumber, region):
        self.userNumber = userNumber
        self._region = region

    def MeshDecompositionSet(self, decomposition):
        self.meshDecomposition = decomposition

    def GeometricFieldSet(self, field):
        self.geometricField = field

    def TypeSet(self, t):
        self.type = t

    def VariableLabelSet(self, vt, label):
        self._var(vt).label = label

    def NumberOfComponentsSet(self, vt, n):
        v = self._var(vt)
        v.ncomp = int(n)
        v.mesh_comp = (v.mesh_comp + [1] * n)[:n]
        v.interp = (v.interp + [v.interp[-1] if v.interp else FieldInterpolationTypes.NODE_BASED] * n)[:n]

    def ComponentMeshComponentSet(self, vt, comp, meshComp):
        v = self._var(vt)
        if comp > v.ncomp:
            self.NumberOfComponentsSet(vt, comp)
        v.mesh_comp[comp - 1] = int(meshComp)

    def ComponentInterpolationSet(self, vt, comp, interp):
        v = self._var(vt)
        if comp > v.ncomp:
            self.NumberOfComponentsSet(vt, comp)
        v.interp[comp - 1] = FieldInterpolationTypes(int(interp))

    def DOFOrderTypeSet(self, vt, order):
        if int(order) != FieldDOFOrderTypes.SEPARATED:
            raise CMFEError(1, "only the SEPARATED DOF order is supported")
        self._dof_order[int(vt)] = int(order)

    def CreateFinish(self):
        self._allocate()
        self._finished = True

    # -- storage
    def _var(self, vt) -> _Variable:
        vt = FieldVariableTypes(int(vt))
        if vt not in self._vars:
            self._vars[vt] = _Variable(self._vars[FieldVariableTypes.U].ncomp)
        return self._vars[vt]

    def _mesh(self):
        return self.meshDecomposition._mesh._regular

    def _allocate(self):
        rm = self._mesh()
        for v in self._vars.values():
            off = [0]
            for c in range(v.ncomp):
                it = v.interp[c]
                if it == FieldInterpolationTypes.NODE_BASED:
                    off.append(off[-1] + rm.num_nodes(v.mesh_comp[c] - 1))
                elif it == FieldInterpolationTypes.ELEMENT_BASED:
                    off.append(off[-1] + rm.num_elements)
                else:
                    off.append(off[-1] + 1)
            v.offsets = off
            v.values = np.zeros(off[-1])

    def _data(self, vt):
        v = self._var(vt)
        if v.values is None:
            self._allocate()
        return v.values

    def _offset(self, vt, comp):
        return self._var(vt).offsets[comp - 1]

    def _component(self, vt, comp):
        v = self._var(vt)
        if v.values is None:
            self._allocate()
        return v.values[v.offsets[comp - 1]: v.offsets[comp]]

    # -- parameter access
    def ComponentValuesInitialiseDP(self, vt, setType, comp, value):
        self._component(vt, comp)[:] = float(value)

    ComponentValuesInitialise = ComponentValuesInitialiseDP

    def ParameterSetUpdateElementDP(self, vt, setType, elementUserNumber, comp, value):
        self._component(vt, comp)[int(elementUserNumber) - 1] = float(value)

    def ParameterSetUpdateNodeDP(self, vt, setType, version, derivative, nodeUserNumber, comp, value):
        rm = self._mesh()
        mc = self._var(vt).mesh_comp[comp - 1] - 1
        self._component(vt, comp)[int(rm.local_node(int(nodeUserNumber), mc))] = float(value)

    def ParameterSetUpdateConstantDP(self, vt, setType, comp, value):
        self._component(vt, comp)[:] = float(value)

    def ParameterSetGetNodeDP(self, vt, setType, version, derivative, nodeUserNumber, comp):
        rm = self._mesh()
        mc = self._var(vt).mesh_comp[comp - 1] - 1
        return float(self._component(vt, comp)[int(rm.local_node(int(nodeUserNumber), mc))])

    def ParameterSetGetElementDP(self, vt, setType, elementUserNumber, comp):
        return float(self._component(vt, comp)[int(elementUserNumber) - 1])

    def ParameterSetDataGet(self, vt, setType):
        """Writable view of all DOF values of the variable (separated order)."""
        return self._data(vt)

    def ParameterSetDataRestore(self, vt, setType, data=None):
        pass

    def ParameterSetUpdateStart(self, vt, setType):
        pass

    def ParameterSetUpdateFinish(self, vt, setType):
        pass

    @staticmethod
    def ParametersToFieldParametersComponentCopy(src, srcVt, srcSet, srcComp, dst, dstVt, dstSet, dstComp):
        dst._component(dstVt, dstComp)[:] = src._component(srcVt, srcComp)


# ------------------------------------------------------- equations sets
class Equations(_Obj):
    def __init__(self):
        self.sparsityType = EquationsSparsityTypes.SPARSE
        self.outputType = EquationsOutputTypes.NONE
        self.lumpingType = 0
        # finite-difference Jacobian step (equations_matrices_routines.F90:356)
        self.jacobianFiniteDifferenceStepSize = 1e-6

    def JacobianFiniteDifferenceStepSizeSet(self, jacobianNumber, stepSize):
        """cmfe_Equations_JacobianFiniteDifferenceStepSizeSet (opencmiss_iron.F90:2325)."""
        if float(stepSize) <= 0.0:
            raise CMFEError(1, "the finite-difference step size must be positive")
        self.jacobianFiniteDifferenceStepSize = float(stepSize)


class EquationsSet(_Obj):
    def __init__(self):
        self.specification = None
        self.geometricField = None
        self.dependentField = None
        self.materialsField = None
        self.sourceField = None
        self.equations = None
        self.reactions = None
        self._region = None

    def CreateStart(self, userNumber, region, geometricField, specification, eqSetFieldUserNumber=None,
                    eqSetField=None):
        self.userNumber = userNumber
        self._region = region
        self.geometricField = geometricField
        self.specification = [int(s) for s in specification]
        cls, typ, sub = (self.specification + [0, 0, 0])[:3]
        ok = ((cls == EquationsSetClasses.CLASSICAL_FIELD and typ == EquationsSetTypes.LAPLACE_EQUATION
               and sub in (1, 2))
              or (cls == EquationsSetClasses.CLASSICAL_FIELD and typ == EquationsSetTypes.POISSON_EQUATION
                  and sub == 1)
              or (cls == EquationsSetClasses.ELASTICITY and typ == EquationsSetTypes.LINEAR_ELASTICITY
                  and sub == 4)
              or (cls == EquationsSetClasses.ELASTICITY and typ == EquationsSetTypes.FINITE_ELASTICITY
                  and sub in (0, 1, 22)))
        if not ok:
            raise CMFEError(1, f"equations set specification {self.specification} is not implemented")

    def CreateFinish(self):
        pass

    @property
    def _kind(self):
        cls, typ, sub = self.specification[:3]
        if cls == EquationsSetClasses.CLASSICAL_FIELD:
            return {(1, 1): "laplace", (1, 2): "generalised_laplace", (2, 1): "poisson"}[(typ, sub)]
        return {1: "linear_elasticity", 2: "mooney_rivlin"}[typ]

    @property
    def _nearly_incompressible(self):
        # no pressure component; materials c1, c2 and the bulk penalty kappa
        return self.specification[2] == EquationsSetSubtypes.NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN

    def _geom_decomp(self):
        return self.geometricField.meshDecomposition

    def _new_field(self, field, ncomp, interp=FieldInterpolationTypes.NODE_BASED, variables=(FieldVariableTypes.U,)):
        field.meshDecomposition = self._geom_decomp()
        field.geometricField = self.geometricField
        field._vars = {}
        for vt in variables:
            v = _Variable(ncomp)
            v.interp = [interp] * ncomp
            field._vars[vt] = v
        return field

    def DependentCreateStart(self, userNumber, field):
        kind = self._kind
        ncomp = {"laplace": 1, "generalised_laplace": 1, "poisson": 1, "linear_elasticity": 3,
                 "mooney_rivlin": 3 if self._nearly_incompressible else 4}[kind]
        self.dependentField = self._new_field(field, ncomp, variables=(FieldVariableTypes.U,
                                                                       FieldVariableTypes.DELUDELN))
        field.type = FieldTypes.GENERAL

    def DependentCreateFinish(self):
        self.dependentField.CreateFinish()

    def MaterialsCreateStart(self, userNumber, field):
        kind = self._kind
        ncomp = {"laplace": 0, "generalised_laplace": 6, "poisson": 1, "linear_elasticity": 6,
                 "mooney_rivlin": 3 if self._nearly_incompressible else 2}[kind]
        variables = (FieldVariableTypes.U, FieldVariableTypes.V) if kind == "mooney_rivlin" else \
            (FieldVariableTypes.U,)
        self.materialsField = self._new_field(field, max(ncomp, 1), FieldInterpolationTypes.CONSTANT, variables)
        if kind == "mooney_rivlin":
            field._vars[FieldVariableTypes.V] = _Variable(1)
            field._vars[FieldVariableTypes.V].interp = [FieldInterpolationTypes.CONSTANT]
        field.type = FieldTypes.MATERIAL

    def MaterialsCreateFinish(self):
        self.materialsField.CreateFinish()
        kind = self._kind
        if kind == "generalised_laplace":
            for c in (1, 2, 3):
                self.materialsField.ComponentValuesInitialiseDP(FieldVariableTypes.U, 1, c, 1.0)
        if kind == "linear_elasticity":  # iron's defaults: E = 10, nu = 0.3 (linear_elasticity_routines.F90)
            for c in (1, 2, 3):
                self.materialsField.ComponentValuesInitialiseDP(FieldVariableTypes.U, 1, c, 10.0)
            for c in (4, 5, 6):
                self.materialsField.ComponentValuesInitialiseDP(FieldVariableTypes.U, 1, c, 0.3)

    def SourceCreateStart(self, userNumber, field):
        ncomp = 3 if self._kind == "mooney_rivlin" else 1
        self.sourceField = self._new_field(field, ncomp, FieldInterpolationTypes.CONSTANT)
        field.type = FieldTypes.GENERAL

    def SourceCreateFinish(self):
        self.sourceField.CreateFinish()

    def EquationsCreateStart(self, equations: Equations):
        self.equations = equations

    def EquationsCreateFinish(self):
        pass


# ----------------------------------------------------- solvers / problem
class ControlLoop(_Obj):
    def __init__(self):
        self.maximumIterations = 1
        self.outputType = 0


class SolverEquations(_Obj):
    def __init__(self):
        self.sparsityType = SolverEquationsSparsityTypes.SPARSE
        self._sets = []
        self._bc = None

    def EquationsSetAdd(self, es):
        self._sets.append(es)
        return len(self._sets)

    def BoundaryConditionsCreateStart(self, bc):
        self._bc = bc
        bc._solver_equations = self

    def BoundaryConditionsCreateFinish(self):
        pass


class BoundaryConditions(_Obj):
    def __init__(self):
        self._entries = []  # (field, vt, node, comp, condition, value, add)
        self._constraints = []  # (field, vt, comp, nodes, coefficient)

    def ConstrainNodeDofsEqual(self, field, fieldVariableType, versionNumber, derivativeNumber, component, nodes,
                               coefficient):
        """BoundaryConditions_ConstrainNodeDofsEqual / ConstrainDofsEqual
        (boundary_condition_routines.F90:3100-3196): the DOF of ``component`` at each of
        nodes[1:] becomes ``coefficient`` times the DOF at nodes[0], one solver DOF for
        all of them (DofConstraintSet, :3204-3313)."""
        nodes = tuple(int(n) for n in nodes)
        if len(nodes) < 2:
            raise CMFEError(1, "Cannot constrain zero or 1 DOF to be equal.")
        if len(set(nodes)) != len(nodes):
            raise CMFEError(1, "a node is duplicated in the DOFs constrained to be equal")
        self._constraints.append((field, int(fieldVariableType), int(component), nodes, float(coefficient)))

    def SetNode(self, field, vt, version, derivative, node, comp, condition, value):
        self._entries.append((field, int(vt), int(node), int(comp), int(condition), float(value), False))

    def AddNode(self, field, vt, version, derivative, node, comp, condition, value):
        self._entries.append((field, int(vt), int(node), int(comp), int(condition), float(value), True))


class Solver(_Obj):
    def __init__(self):
        self.outputType = SolverOutputTypes.NONE
        self.linearType = LinearSolverTypes.ITERATIVE
        # KSP defaults, solver_routines.F90:11320-11369
        self.linearIterativeType = IterativeLinearSolverTypes.GMRES
        self.linearIterativePreconditionerType = IterativePreconditionerTypes.JACOBI
        self.linearIterativeRelativeTolerance = 1.0e-5
        self.linearIterativeAbsoluteTolerance = 1.0e-10
        self.linearIterativeDivergenceTolerance = 1.0e5
        self.linearIterativeMaximumIterations = 100000
        self.linearIterativeGMRESRestart = 30
        self.linearIterativePrecision = "fp64"
        # Newton defaults, solver_routines.F90:17960-18030
        self.newtonMaximumIterations = 50
        self.newtonMaximumFunctionEvaluations = 1000
        self.newtonRelativeTolerance = 1.0e-5
        self.newtonAbsoluteTolerance = 1.0e-10
        self.newtonSolutionTolerance = 1.0e-5
        self.newtonJacobianCalculationType = JacobianCalculationTypes.FD
        self.newtonLineSearchType = NewtonLineSearchTypes.CUBIC
        self.newtonConvergenceTestType = NewtonConvergenceTypes.PETSC_DEFAULT
        self.newtonLineSearchAlpha = 1e-4
        self._problem = None
        self._linear = None
        self._solver_equations = None
        self.last_info = None

    def SolverEquationsGet(self, se: SolverEquations):
        owner = self._problem._solver_equations if self._problem is not None else None
        if owner is None:
            raise CMFEError(1, "solver equations have not been created")
        se.__dict__ = owner.__dict__

    def NewtonLinearSolverGet(self, solver):
        if self._linear is None:
            self._linear = Solver()
        solver.__dict__ = self._linear.__dict__

    def NewtonJacobianCalculationTypeSet(self, t):
        self.newtonJacobianCalculationType = JacobianCalculationTypes(int(t))

    def NewtonConvergenceTestTypeSet(self, t):
        """cmfe_Solver_NewtonConvergenceTestTypeSet (solver_routines.F90:18349): PETSC_DEFAULT
        or ENERGY_NORM (DIFFERENTIATED_RATIO raises at solve time, as in iron)."""
        self.newtonConvergenceTestType = NewtonConvergenceTypes(int(t))

    def NewtonLineSearchTypeSet(self, t):
        """cmfe_Solver_NewtonLineSearchTypeSet: NONORMS, LINEAR, QUADRATIC or CUBIC."""
        self.newtonLineSearchType = NewtonLineSearchTypes(int(t))

    def NewtonLineSearchAlphaSet(self, alpha):
        """cmfe_Solver_NewtonLineSearchAlphaSet: sufficient-decrease parameter."""
        if float(alpha) <= 0.0:
            raise CMFEError(1, "the line search alpha must be positive")
        self.newtonLineSearchAlpha = float(alpha)


class Problem(_Obj):
    def __init__(self):
        self.specification = None
        self._solver = Solver()
        self._solver._problem = self
        self._control_loop = ControlLoop()
        self._solver_equations = None

    def CreateStart(self, userNumber, specification):
        self.specification = [int(s) for s in specification]

    def CreateFinish(self):
        pass

    def ControlLoopCreateStart(self):
        pass

    def ControlLoopGet(self, identifiers, loop: ControlLoop):
        loop.__dict__ = self._control_loop.__dict__

    def ControlLoopCreateFinish(self):
        pass

    def SolversCreateStart(self):
        pass

    def SolverGet(self, identifiers, index, solver: Solver):
        solver.__dict__ = self._solver.__dict__

    def SolversCreateFinish(self):
        pass

    def SolverEquationsCreateStart(self):
        self._solver_equations = SolverEquations()

    def SolverEquationsCreateFinish(self):
        pass

    def Solve(self):
        from . import engine
        se = self._solver_equations
        if se is None or not se._sets:
            raise CMFEError(1, "no equations set was added to the solver equations")
        bc = se._bc or BoundaryConditions()
        if len(se._sets) > 1:
            # several sets in one solver system (no interface conditions: block diagonal)
            if any(es._kind == "mooney_rivlin" for es in se._sets):
                raise CMFEError(1, "several equations sets per solver equations are implemented for linear sets")
            engine.solve_linear_sets(se._sets, self._solver, bc)
            return
        engine.solve(self, se._sets[0], self._solver, bc)
