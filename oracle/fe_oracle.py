"""FP64 CPU restatement of OpenCMISS-Iron's element-assembly and sparse-solve path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: tests/, the smoke check and
bench.py's cpu_baseline leg use it as the checker.  The product package (iron_amd/)
never imports it, and nothing here is ever the thing that is measured or shipped.

Each function restates the algorithm of the reference file:line named in its
docstring (paths relative to the reference's root).  Everything is float64 and
1-based numbering is converted to 0-based only at the array boundary.

Parity pinning (what checks each part; see tests/test_oracle_*.py):
  * gauss_legendre / basis tables   -> exact-integration and partition-of-unity
    identities, the closed-form tables of basis_routines.F90:6822-6835.
  * regular_mesh / dof numbering / csr_pattern -> brute-force enumeration on tiny
    meshes, nnz closed forms (1-D factor products), node coordinates of the
    reference's committed .exnode files.
  * laplace / poisson / elasticity element matrices -> closed forms on boxes
    (h*K_ref scaling, row sums, rigid body modes, lambda/mu textbook tensor).
  * assembly + Dirichlet lifting + solve + back-substitution -> the reference's
    golden AnalyticLaplaceLinearLagrange.opanal (solution and the iron-compat
    "del Phi/del n" values) and the closed form u = x of the unit-cube problem.
  * constraint_couplings / reduce_coupled / field_update_coupled -> equal to
    reduce_dirichlet without constraints, to a dense P^T K P, to the dense KKT
    (Lagrange multiplier) solve of the constrained problem, and to the closed form
    u = x with periodic faces; iron's constraint error rules.
  * Mooney-Rivlin residual/Jacobian -> zero residual at the reference state,
    central finite differences of the residual, and the homogeneous uniaxial
    stretch closed form.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# constants.F90:69-70
ZERO_TOLERANCE = 5.0 * np.finfo(np.float64).eps

# Partial derivative indices, constants.F90:176-188 (1-based in the reference).
NO_PART_DERIV = 1
PART_DERIV_S1 = 2
PART_DERIV_S2 = 4
PART_DERIV_S3 = 7
FIRST_DERIV_MAP = (PART_DERIV_S1, PART_DERIV_S2, PART_DERIV_S3)  # constants.F90:253
# PARTIAL_DERIVATIVE_INDEX(nu, xi) -> derivative order per xi direction
# (1 = value, 2 = first, 3 = second), constants.F90:236-251, for nu = 1..11.
_PD_INDEX = {
    1: (1, 1, 1), 2: (2, 1, 1), 3: (3, 1, 1), 4: (1, 2, 1), 5: (1, 3, 1),
    6: (2, 2, 1), 7: (1, 1, 2), 8: (1, 1, 3), 9: (2, 1, 2), 10: (1, 2, 2),
    11: (2, 2, 2),
}

LINEAR, QUADRATIC, CUBIC = 1, 2, 3  # polynomial order of a Lagrange basis


# ----------------------------------------------------------------------------
# a4: quadrature and basis tabulation
# ----------------------------------------------------------------------------
def gauss_legendre(n: int):
    """Gauss-Legendre points/weights on [0,1]; basis_routines.F90:6804-6893."""
    s6 = math.sqrt(6.0 / 5.0)
    loc = [0.5,
           (-1.0 / math.sqrt(3.0) + 1.0) / 2.0, (+1.0 / math.sqrt(3.0) + 1.0) / 2.0,
           (-math.sqrt(0.6) + 1.0) / 2.0, 0.5, (+math.sqrt(0.6) + 1.0) / 2.0,
           (-math.sqrt((3.0 + 2.0 * s6) / 7.0) + 1.0) / 2.0,
           (-math.sqrt((3.0 - 2.0 * s6) / 7.0) + 1.0) / 2.0,
           (+math.sqrt((3.0 - 2.0 * s6) / 7.0) + 1.0) / 2.0,
           (+math.sqrt((3.0 + 2.0 * s6) / 7.0) + 1.0) / 2.0]
    wts = [1.0, 0.5, 0.5, 2.5 / 9.0, 4.0 / 9.0, 2.5 / 9.0,
           (18.0 - math.sqrt(30.0)) / 72.0, (18.0 + math.sqrt(30.0)) / 72.0,
           (18.0 + math.sqrt(30.0)) / 72.0, (18.0 - math.sqrt(30.0)) / 72.0]
    if n < 1 or n > 4:
        raise ValueError(f"The specified number of Gauss points of {n} is invalid or not implemented.")
    start = [0, 1, 3, 6][n - 1]
    return np.array(loc[start:start + n]), np.array(wts[start:start + n])


def lagrange_1d(order: int, node: int, pd: int, xi: float) -> float:
    """1-D Lagrange functions; basis_routines.F90:7969-8091 (linear, quadratic),
    :7899 (cubic).  ``node`` is 1-based, ``pd`` is 1 (value), 2 (d/dxi), 3 (d2/dxi2)."""
    if order == LINEAR:
        if pd == 1:
            return (1.0 - xi) if node == 1 else xi
        if pd == 2:
            return -1.0 if node == 1 else 1.0
        return 0.0
    if order == QUADRATIC:
        if pd == 1:
            return [1.0 - 3.0 * xi + 2.0 * xi * xi, 4.0 * xi * (1.0 - xi), xi * (xi + xi - 1.0)][node - 1]
        if pd == 2:
            return [4.0 * xi - 3.0, 4.0 - 8.0 * xi, 4.0 * xi - 1.0][node - 1]
        return [4.0, -8.0, 4.0][node - 1]
    if order == CUBIC:
        # Lagrange cubic through 0, 1/3, 2/3, 1 (basis_routines.F90:7899-7966)
        if pd == 1:
            return [0.5 * (3.0 * xi - 1.0) * (3.0 * xi - 2.0) * (1.0 - xi),
                    4.5 * xi * (3.0 * xi - 2.0) * (xi - 1.0),
                    4.5 * xi * (3.0 * xi - 1.0) * (1.0 - xi),
                    0.5 * xi * (3.0 * xi - 1.0) * (3.0 * xi - 2.0)][node - 1]
        if pd == 2:
            return [-13.5 * xi * xi + 18.0 * xi - 5.5,
                    40.5 * xi * xi - 45.0 * xi + 9.0,
                    -40.5 * xi * xi + 36.0 * xi - 4.5,
                    13.5 * xi * xi - 9.0 * xi + 1.0][node - 1]
        return [18.0 - 27.0 * xi, -45.0 + 81.0 * xi, 36.0 - 81.0 * xi, -9.0 + 27.0 * xi][node - 1]
    raise ValueError("invalid Lagrange order")


@dataclass
class Basis:
    """A Lagrange tensor-product basis plus its default Gauss scheme.

    Local node order is xi1-fastest lexicographic (basis_routines.F90:2551-2590);
    Gauss point order is xi1-fastest, ng = i + (j-1 + (k-1) n2) n1
    (basis_routines.F90:3813-3841).  ``gauss_basis_fns[ns, nu, ng]`` follows
    GAUSS_BASIS_FNS(ns,nu,ng) with 0-based ns/ng and nu indexed 1..11 (slot 0 unused).
    """
    nxi: int
    order: int
    ngauss_xi: int
    nodes_xic: int = 0
    node_index: np.ndarray = None      # [nn, nxi] 1-based per-direction node index
    gauss_positions: np.ndarray = None  # [ng, nxi]
    gauss_weights: np.ndarray = None    # [ng]
    gauss_basis_fns: np.ndarray = None  # [nn, 12, ng]

    def __post_init__(self):
        self.nodes_xic = self.order + 1
        n = self.nodes_xic
        idx = []
        for k in range(n if self.nxi > 2 else 1):
            for j in range(n if self.nxi > 1 else 1):
                for i in range(n):
                    idx.append((i + 1, j + 1, k + 1)[: self.nxi])
        self.node_index = np.array(idx, dtype=np.int64)
        x1, w1 = gauss_legendre(self.ngauss_xi)
        g = self.ngauss_xi
        pos, wts = [], []
        for k in range(g if self.nxi > 2 else 1):
            for j in range(g if self.nxi > 1 else 1):
                for i in range(g):
                    p = (x1[i], x1[j], x1[k])[: self.nxi]
                    w = w1[i] * (w1[j] if self.nxi > 1 else 1.0) * (w1[k] if self.nxi > 2 else 1.0)
                    pos.append(p)
                    wts.append(w)
        self.gauss_positions = np.array(pos)
        self.gauss_weights = np.array(wts)
        nn = len(idx)
        ng = len(pos)
        tab = np.zeros((nn, 12, ng))
        for gi in range(ng):
            for a in range(nn):
                for nu in range(1, 12):
                    pdi = _PD_INDEX[nu]
                    if self.nxi < 3 and pdi[2] != 1:
                        continue
                    if self.nxi < 2 and pdi[1] != 1:
                        continue
                    tab[a, nu, gi] = self.evaluate(a, nu, self.gauss_positions[gi])
        self.gauss_basis_fns = tab

    @property
    def number_of_nodes(self) -> int:
        return len(self.node_index)

    def evaluate(self, a: int, nu: int, xi) -> float:
        """BASIS_LHTP_BASIS_EVALUATE_DP, basis_routines.F90:3243 (0-based node a)."""
        v = 1.0
        pdi = _PD_INDEX[nu]
        for d in range(self.nxi):
            v *= lagrange_1d(self.order, int(self.node_index[a, d]), pdi[d], float(xi[d]))
        return v

    def dxi_table(self) -> np.ndarray:
        """[ng, nn, nxi] first derivatives dphi/dxi at the Gauss points."""
        return np.stack([self.gauss_basis_fns[:, FIRST_DERIV_MAP[d], :].T for d in range(self.nxi)], axis=-1)

    def value_table(self) -> np.ndarray:
        """[ng, nn] basis values at the Gauss points."""
        return self.gauss_basis_fns[:, NO_PART_DERIV, :].T.copy()


# ----------------------------------------------------------------------------
# a1-a3: regular generated mesh, mesh nodes, field DOF numbering
# ----------------------------------------------------------------------------
@dataclass
class RegularMesh:
    """GENERATED_MESH_REGULAR_CREATE_FINISH, generated_mesh_routines.F90:1389-2214.

    ``elem_nodes[b]`` holds user node numbers (1-based) of each element for mesh
    component / basis ``b``; elements are numbered xi1-fastest
    (generated_mesh_routines.F90:1568-1612).  ``comp_nodes[b]`` are the sorted
    unique user numbers of mesh component ``b`` (MeshTopology_NodesCalculate,
    mesh_routines.F90:9159-9279); local node number = position in that list
    (one domain: local == global order, mesh_routines.F90:4603-4620).
    """
    n_elem: tuple
    bases: list
    extent: tuple
    origin: tuple
    elem_nodes: list = field(default_factory=list)
    comp_nodes: list = field(default_factory=list)
    coords: np.ndarray = None  # [N1, dim] coordinates of mesh-component-1 nodes (by local index)

    @property
    def dim(self):
        return len(self.n_elem)

    @property
    def num_elements(self):
        return int(np.prod(self.n_elem))


def regular_mesh(n_elem, bases, extent, origin=None) -> RegularMesh:
    dim = len(n_elem)
    if origin is None:
        origin = (0.0,) * dim
    m = RegularMesh(tuple(n_elem), list(bases), tuple(extent), tuple(origin))
    E = m.num_elements
    b0 = bases[0]
    for bi, basis in enumerate(bases):
        nx = basis.nodes_xic
        tot = [(nx - 1) * n_elem[d] + 1 for d in range(dim)]
        en = np.zeros((E, basis.number_of_nodes), dtype=np.int64)
        for ne3 in range(1, (n_elem[2] if dim > 2 else 1) + 1):
            for ne2 in range(1, (n_elem[1] if dim > 1 else 1) + 1):
                for ne1 in range(1, n_elem[0] + 1):
                    grid_ne = ne1
                    np0 = 1 + (ne1 - 1) * (nx - 1)
                    if dim > 1:
                        grid_ne += (ne2 - 1) * n_elem[0]
                        np0 += (ne2 - 1) * tot[0] * (nx - 1)
                        if dim > 2:
                            grid_ne += (ne3 - 1) * n_elem[0] * n_elem[1]
                            np0 += (ne3 - 1) * tot[0] * tot[1] * (nx - 1)
                    comp = []
                    for a in range(basis.number_of_nodes):
                        ii = basis.node_index[a]
                        off = ii[0] - 1
                        if dim > 1:
                            off += (ii[1] - 1) * tot[0]
                        if dim > 2:
                            off += (ii[2] - 1) * tot[0] * tot[1]
                        comp.append(np0 + off)
                    if bi == 0:
                        en[grid_ne - 1] = comp
                    else:
                        # GeneratedMesh_RegularComponentNodesToUserNumbers,
                        # generated_mesh_routines.F90:4941-5357: a coarser basis takes
                        # the user numbers of the first basis' corner nodes.
                        if nx == b0.nodes_xic:
                            en[grid_ne - 1] = m.elem_nodes[0][grid_ne - 1]
                        elif nx == 2:
                            for a in range(basis.number_of_nodes):
                                ii = basis.node_index[a]
                                fi = [1 if ii[d] == 1 else b0.nodes_xic for d in range(dim)]
                                a0 = fi[0] - 1
                                if dim > 1:
                                    a0 += (fi[1] - 1) * b0.nodes_xic
                                if dim > 2:
                                    a0 += (fi[2] - 1) * b0.nodes_xic * b0.nodes_xic
                                en[grid_ne - 1, a] = m.elem_nodes[0][grid_ne - 1, a0]
                        else:
                            raise NotImplementedError("only linear secondary bases are restated")
        m.elem_nodes.append(en)
        m.comp_nodes.append(np.unique(en))
    # GeneratedMesh_RegularGeometricParametersCalculate, generated_mesh_routines.F90:3217-3384
    nx = b0.nodes_xic
    tot = [(nx - 2) * n_elem[d] + n_elem[d] + 1 for d in range(dim)]
    delta = [extent[d] / float(tot[d] - 1) for d in range(dim)]
    user = m.comp_nodes[0]
    cn = user - 1
    pos = []
    rem = cn.copy()
    stride = 1
    strides = []
    for d in range(dim):
        strides.append(stride)
        stride *= tot[d]
    for d in reversed(range(dim)):
        pos.insert(0, rem // strides[d])
        rem = rem % strides[d]
    coords = np.zeros((len(user), dim))
    for c in range(dim):
        v = np.zeros(len(user))
        for d in range(dim):
            dc = delta[d] if d == c else 0.0
            v = v + pos[d].astype(np.float64) * dc
        coords[:, c] = origin[c] + v
    m.coords = coords
    return m


@dataclass
class DofMap:
    """FIELD_MAPPINGS_CALCULATE with SEPARATED DOF order, field_routines.F90:9883-11116
    (separated branch :10056): all DOFs of component 1 (node-minor), then component 2, ...

    ``comp_mesh[c]`` is the mesh component of field component ``c``.  Node-based
    Lagrange DOFs only (one derivative, one version per node).
    """
    mesh: RegularMesh
    comp_mesh: list
    offsets: list = field(default_factory=list)
    num_dofs: int = 0
    node_local: list = field(default_factory=list)  # user number -> local index per mesh comp

    def __post_init__(self):
        off = 0
        for c, mc in enumerate(self.comp_mesh):
            self.offsets.append(off)
            off += len(self.mesh.comp_nodes[mc])
        self.num_dofs = off
        for nodes in self.mesh.comp_nodes:
            lut = np.full(int(nodes.max()) + 1, -1, dtype=np.int64)
            lut[nodes] = np.arange(len(nodes))
            self.node_local.append(lut)

    def dof(self, c: int, user_node) -> np.ndarray:
        """0-based DOF of component c (0-based) at user node number(s)."""
        mc = self.comp_mesh[c]
        return self.offsets[c] + self.node_local[mc][np.asarray(user_node)]

    def element_dofs(self) -> np.ndarray:
        """Element DOF list [E, n_edofs]: component-major then element node
        (EquationsMatrices_ElementMatrixCalculate, equations_matrices_routines.F90:609-840,
        rows = local DOF, columns = global DOF, :674-680)."""
        cols = []
        for c, mc in enumerate(self.comp_mesh):
            cols.append(self.dof(c, self.mesh.elem_nodes[mc]))
        return np.concatenate(cols, axis=1)


# ----------------------------------------------------------------------------
# a14: CSR sparsity (FEM structure) of the equations matrix
# ----------------------------------------------------------------------------
def csr_pattern(dm: DofMap):
    """EquationsMatrix_StructureCalculate (FEM, CSR), equations_matrices_routines.F90:4399-4461.

    For each row DOF: union of every element DOF (all components) over the elements
    surrounding the row's node, then LIST_REMOVE_DUPLICATES (sort + unique,
    lists.F90:2652-2700).  Returns 0-based (row_ptr[int64], col[int64]).
    """
    edofs = dm.element_dofs()
    mesh = dm.mesh
    rows = []
    for c, mc in enumerate(dm.comp_mesh):
        en = mesh.elem_nodes[mc]
        nodes = mesh.comp_nodes[mc]
        # surrounding elements of each node of this mesh component
        surround = [[] for _ in range(len(nodes))]
        loc = dm.node_local[mc]
        for e in range(en.shape[0]):
            for a in range(en.shape[1]):
                surround[loc[en[e, a]]].append(e)
        for li in range(len(nodes)):
            els = sorted(set(surround[li]))
            cand = edofs[els].ravel() if els else np.zeros(0, dtype=np.int64)
            rows.append(np.unique(cand))
    row_ptr = np.zeros(len(rows) + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    return row_ptr, col


def storage_location_find(row_ptr, col, i, j) -> int:
    """MATRIX_STORAGE_LOCATION_FIND for CSR, matrix_vector.F90:1724-1854: bisection
    until at most bisectionToLinearSearchThreshold (=10, :184) entries remain, then a
    linear scan.  Returns the 0-based position, or -1 if (i,j) is not in the pattern."""
    lo = int(row_ptr[i])
    hi = int(row_ptr[i + 1]) - 1
    while hi - lo + 1 > 10:
        mid = (lo + hi) // 2
        if col[mid] == j:
            return mid
        if col[mid] < j:
            lo = mid + 1
        else:
            hi = mid - 1
    for k in range(lo, hi + 1):
        if col[k] == j:
            return k
    return -1


# ----------------------------------------------------------------------------
# a5-a9: element routines
# ----------------------------------------------------------------------------
def _invert_full(A):
    """InvertFullDP, maths.F90:2807-2868 (batched over leading axes, 2x2 or 3x3)."""
    n = A.shape[-1]
    if n == 2:
        det = A[..., 0, 0] * A[..., 1, 1] - A[..., 1, 0] * A[..., 0, 1]
        B = np.empty_like(A)
        B[..., 0, 0] = A[..., 1, 1] / det
        B[..., 0, 1] = -A[..., 0, 1] / det
        B[..., 1, 0] = -A[..., 1, 0] / det
        B[..., 1, 1] = A[..., 0, 0] / det
        return B, det
    a = lambda i, j: A[..., i - 1, j - 1]
    det = (a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2)
           - a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3) - a(3, 1) * a(2, 2) * a(1, 3))
    B = np.empty_like(A)
    B[..., 0, 0] = (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) / det
    B[..., 1, 0] = (a(2, 3) * a(3, 1) - a(3, 3) * a(2, 1)) / det
    B[..., 2, 0] = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) / det
    B[..., 0, 1] = (a(3, 2) * a(1, 3) - a(1, 2) * a(3, 3)) / det
    B[..., 1, 1] = (a(3, 3) * a(1, 1) - a(1, 3) * a(3, 1)) / det
    B[..., 2, 1] = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) / det
    B[..., 0, 2] = (a(1, 2) * a(2, 3) - a(2, 2) * a(1, 3)) / det
    B[..., 1, 2] = (a(1, 3) * a(2, 1) - a(2, 3) * a(1, 1)) / det
    B[..., 2, 2] = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) / det
    return B, det


def gauss_metrics(xe, basis: Basis):
    """FIELD_INTERPOLATE_GAUSS (field_routines.F90:5771-5982) of the geometric field and
    COORDINATE_METRICS_CALCULATE (coordinate_routines.F90:639-1026), rectangular
    cartesian, volume Jacobian type.

    xe: [E, nn, dim] element nodal coordinates (FIELD_INTERPOLATION_PARAMETERS_ELEMENT_GET).
    Returns dx_dxi [E,ng,dim,dim] (DX_DXI(i,j)=dx_i/dxi_j), gu [E,ng,dim,dim] (GL^-1),
    jac [E,ng] (=sqrt|det GL|), dxi_dx [E,ng,dim,dim] (DX_DXI^-1, DXI_DX(i,j)=dxi_i/dx_j).
    """
    dN = basis.dxi_table()  # [ng, nn, nxi]
    dx_dxi = np.einsum("ead,gaj->egdj", xe, dN)
    gl = np.einsum("egki,egkj->egij", dx_dxi, dx_dxi)
    gu, det_gl = _invert_full(gl)
    jac = np.sqrt(np.abs(det_gl))
    dxi_dx, _ = _invert_full(dx_dxi)
    return dx_dxi, gu, jac, dxi_dx


def laplace_element_matrices(xe, basis: Basis, kappa=None):
    """LaplaceEquation_FiniteElementCalculate, Laplace_equations_routines.F90:482-579
    (standard) and :581-747 (generalised, isotropic conductivity kappa*I, no fibres:
    the xi-space conductivity dNudXi.K.dXidNu reduces to kappa*I).

    K(ms,ns) += sum_{mi,ni} dphi_ms/dxi_mi * dphi_ns/dxi_ni * GU(mi,ni) * |J| * w_g.
    Returns [E, nn, nn] FP64 (row ms, column ns).
    """
    dN = basis.dxi_table()
    _, gu, jac, _ = gauss_metrics(xe, basis)
    rwg = jac * basis.gauss_weights[None, :]
    K = np.einsum("gmi,egij,gnj,eg->emn", dN, gu, dN, rwg)
    if kappa is not None:
        K = K * np.asarray(kappa, dtype=np.float64)[:, None, None]
    return K


def poisson_source_vectors(xe, basis: Basis, source):
    """Constant-source Poisson source vector, Poisson_equations_routines.F90:3407-3415:
    s(ms) += f * phi_ms(g) * |J| * w_g; ``source`` is f per element (materials comp 1)."""
    phi = basis.value_table()
    _, _, jac, _ = gauss_metrics(xe, basis)
    rwg = jac * basis.gauss_weights[None, :]
    f = np.broadcast_to(np.asarray(source, dtype=np.float64), (xe.shape[0],))
    return np.einsum("gm,eg->em", phi, rwg) * f[:, None]


def linear_elasticity_tensor(E1, E2, E3, v13, v23, v12):
    """LINEAR_ELASTICITY_TENSOR (3-D subtype), linear_elasticity_routines.F90:1410-1451.
    Inputs are arrays per element; returns C [E, 6, 6] (Voigt, iron's index use)."""
    E1, E2, E3, v13, v23, v12 = (np.asarray(a, dtype=np.float64) for a in (E1, E2, E3, v13, v23, v12))
    v31, v32, v21 = v13, v23, v12
    gama = 1.0 / (1.0 - v12 * v21 - v23 * v32 - v31 * v13 - 2.0 * v21 * v32 * v13)
    C11 = E1 * (1.0 - v23 * v32) * gama
    C22 = E2 * (1.0 - v13 * v31) * gama
    C33 = E3 * (1.0 - v12 * v21) * gama
    C12 = E1 * (v21 + v31 * v23) * gama
    C13 = E1 * (v31 + v21 * v32) * gama
    C23 = E2 * (v32 + v12 * v31) * gama
    C44 = E2 / (2.0 * (1.0 + v23))
    C55 = E1 / (2.0 * (1.0 + v13))
    C66 = E3 / (2.0 * (1.0 + v12))
    n = np.broadcast(E1, E2).shape
    C = np.zeros(n + (6, 6))
    C[..., 0, 0], C[..., 1, 0], C[..., 2, 0] = C11, C12, C13
    C[..., 0, 1], C[..., 1, 1], C[..., 2, 1] = C12, C22, C23
    C[..., 0, 2], C[..., 1, 2], C[..., 2, 2] = C13, C23, C33
    C[..., 3, 3], C[..., 4, 4], C[..., 5, 5] = C44, C55, C66
    return C


def linear_elasticity_element_matrices(xe, basis: Basis, C):
    """LINEAR_ELASTICITY_FINITE_ELEMENT_CALCULATE, 3-D subtype,
    linear_elasticity_routines.F90:1109-1403.  Element DOFs are component-blocked
    [u1(1..nn), u2(1..nn), u3(1..nn)].

    dphidx(ns,mi) = sum_ni DXI_DX(mi,ni) * dphi_ns/dxi_ni  (:1245-1252, restated as
    written; on axis-aligned boxes DXI_DX is diagonal).  Diagonal blocks use
    JRWG*[C11,C66,C44 | C66,C22,C55 | C44,C55,C33] (:1255-1275), off-diagonal blocks
    (1,2),(1,3),(2,3) use [C12,C66], [C13,C44], [C23,C55] (:1256-1286); the upper
    triangle is mirrored to the lower (:1373-1380).
    """
    E = xe.shape[0]
    nn = basis.number_of_nodes
    dN = basis.dxi_table()
    _, _, jac, dxi_dx = gauss_metrics(xe, basis)
    jrwg = jac * basis.gauss_weights[None, :]                       # [E, ng]
    dphidx = np.einsum("egmn,gsn->egsm", dxi_dx, dN)               # [E, ng, nn, 3]
    C = np.asarray(C)
    Cg = lambda i, j: C[:, i - 1, j - 1][:, None] * jrwg           # [E, ng]
    diag = [[Cg(1, 1), Cg(6, 6), Cg(4, 4)],
            [Cg(6, 6), Cg(2, 2), Cg(5, 5)],
            [Cg(4, 4), Cg(5, 5), Cg(3, 3)]]
    K = np.zeros((E, 3 * nn, 3 * nn))
    for xi in range(3):
        blk = np.zeros((E, nn, nn))
        for d in range(3):
            blk += np.einsum("egn,egm,eg->enm", dphidx[..., d], dphidx[..., d], diag[xi][d])
        K[:, xi * nn:(xi + 1) * nn, xi * nn:(xi + 1) * nn] = blk
    offs = [((0, 1), Cg(1, 2), Cg(6, 6)), ((0, 2), Cg(1, 3), Cg(4, 4)), ((1, 2), Cg(2, 3), Cg(5, 5))]
    for (r, c), ca, cb in offs:
        blk = (np.einsum("egn,egm,eg->enm", dphidx[..., r], dphidx[..., c], ca)
               + np.einsum("egn,egm,eg->enm", dphidx[..., c], dphidx[..., r], cb))
        K[:, r * nn:(r + 1) * nn, c * nn:(c + 1) * nn] = blk
        K[:, c * nn:(c + 1) * nn, r * nn:(r + 1) * nn] = np.transpose(blk, (0, 2, 1))
    return K


# ----------------------------------------------------------------------------
# a13: element add (scatter) into the CSR values
# ----------------------------------------------------------------------------
def assemble_csr(row_ptr, col, edofs, Ke, n_rows):
    """EquationsMatrices_ElementAdd -> MATRIX_VALUES_ADD_DP2 (CSR),
    equations_matrices_routines.F90:1140-1265, matrix_vector.F90:3119-3350.
    Element-by-element in element order, each entry located in its row (the
    located position is independent of the search method, so a vectorised sorted
    search stands in for storage_location_find)."""
    vals = np.zeros(len(col))
    E, n = edofs.shape
    r = np.repeat(edofs, n, axis=1).ravel()
    c = np.tile(edofs, (1, n)).ravel()
    key_pat = np.repeat(np.arange(n_rows), np.diff(row_ptr)) * np.int64(n_rows + 1) + col
    key = r * np.int64(n_rows + 1) + c
    pos = np.searchsorted(key_pat, key)
    if np.any(pos >= len(col)) or np.any(key_pat[np.minimum(pos, len(col) - 1)] != key):
        raise ValueError("element entry outside the CSR pattern")
    np.add.at(vals, pos, Ke.reshape(E, -1).ravel())
    return vals


def assemble_vector(edofs, fe, n_rows):
    v = np.zeros(n_rows)
    np.add.at(v, edofs.ravel(), fe.ravel())
    return v


# ----------------------------------------------------------------------------
# a15-a16: solver mapping, Dirichlet reduction and RHS lifting
# ----------------------------------------------------------------------------
def reduce_dirichlet(row_ptr, col, vals, fixed_mask, u_fixed, rhs_extra, rhs_fixed_mask=None):
    """SOLVER_MAPPING_CALCULATE (solver_mapping_routines.F90:422-547, :1445-1511),
    SOLVER_MATRIX_STRUCTURE_CALCULATE (solver_matrices_routines.F90:1612-2431),
    SOLVER_MATRIX_EQUATIONS_MATRIX_ADD (:806-1023) and the RHS of
    SOLVER_MATRICES_STATIC_ASSEMBLE (solver_routines.F90:14640-14850).

    Solver rows/columns are the free DOFs in ascending order.  b_f = rhs_extra_f
    - sum_d K_fd u_d, where a fixed column contributes only if |u_d| >= ZERO_TOLERANCE
    and |K_fd| >= ZERO_TOLERANCE, and only on rows whose RHS-variable DOF is free (the
    lifting sits in the BOUNDARY_CONDITION_DOF_FREE branch, :14681-14849).
    Returns (free, s_row_ptr, s_col, s_vals, b).
    """
    n = len(fixed_mask)
    free = np.nonzero(~fixed_mask)[0]
    new_idx = np.full(n, -1, dtype=np.int64)
    new_idx[free] = np.arange(len(free))
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    keep = (~fixed_mask[rows]) & (~fixed_mask[col])
    s_rows = new_idx[rows[keep]]
    s_col = new_idx[col[keep]]
    s_vals = vals[keep]
    s_row_ptr = np.zeros(len(free) + 1, dtype=np.int64)
    np.add.at(s_row_ptr, s_rows + 1, 1)
    s_row_ptr = np.cumsum(s_row_ptr)
    b = np.asarray(rhs_extra, dtype=np.float64)[free].copy()
    lift = (~fixed_mask[rows]) & fixed_mask[col]
    if rhs_fixed_mask is not None:
        lift &= ~np.asarray(rhs_fixed_mask, dtype=bool)[rows]
    ud = np.asarray(u_fixed, dtype=np.float64)[col[lift]]
    kv = vals[lift]
    ok = (np.abs(ud) >= ZERO_TOLERANCE) & (np.abs(kv) >= ZERO_TOLERANCE)
    np.add.at(b, new_idx[rows[lift][ok]], -1.0 * kv[ok] * ud[ok])
    return free, s_row_ptr, s_col, s_vals, b


# ----------------------------------------------------------------------------
# a17: Krylov solvers with Jacobi preconditioning (PETSc KSPCG / KSPGMRES + PCJACOBI)
# ----------------------------------------------------------------------------
def csr_matvec(row_ptr, col, vals, x):
    rows = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    y = np.zeros(len(row_ptr) - 1)
    np.add.at(y, rows, vals * x[col])
    return y


def _csr_scipy(row_ptr, col, vals):
    import scipy.sparse as sp
    n = len(row_ptr) - 1
    return sp.csr_matrix((vals, col, row_ptr), shape=(n, n))


def cg_jacobi(row_ptr, col, vals, b, x0=None, rtol=1e-10, maxit=100000):
    """Preconditioned CG (PETSc KSPCG semantics with PCJACOBI, the solver set up by
    SOLVER_LINEAR_ITERATIVE_CREATE_FINISH, solver_routines.F90:10921-11157), stopped on
    the true relative residual ||b - A x|| / ||b|| <= rtol (the stopping rule of this
    build; PETSc's preconditioned-norm test is not restated).  Returns (x, iters, relres)."""
    A = _csr_scipy(row_ptr, col, vals)
    d = A.diagonal()
    x = np.zeros_like(b) if x0 is None else x0.astype(np.float64).copy()
    r = b - A @ x
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, 0.0
    z = r / d
    p = z.copy()
    rz = r @ z
    it = 0
    while it < maxit:
        if np.linalg.norm(r) / bn <= rtol:
            break
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = r / d
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return x, it, float(np.linalg.norm(b - A @ x) / bn)


def gmres_jacobi(row_ptr, col, vals, b, x0=None, rtol=1e-10, restart=30, maxit=100000):
    """Right-Jacobi-preconditioned restarted GMRES(30) (PETSc KSPGMRES default restart,
    solver_routines.F90:11320-11369), modified Gram-Schmidt Arnoldi with Givens
    rotations, stopped on the true relative residual."""
    A = _csr_scipy(row_ptr, col, vals)
    d = A.diagonal()
    x = np.zeros_like(b) if x0 is None else x0.astype(np.float64).copy()
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, 0.0
    it = 0
    while it < maxit:
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta / bn <= rtol:
            break
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k_used = 0
        for k in range(restart):
            w = A @ (V[k] / d)
            for j in range(k + 1):
                H[j, k] = w @ V[j]
                w = w - H[j, k] * V[j]
            H[k + 1, k] = np.linalg.norm(w)
            V.append(w / H[k + 1, k] if H[k + 1, k] != 0.0 else w)
            for j in range(k):
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
            H[k, k] = den
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            k_used = k + 1
            if abs(g[k + 1]) / bn <= rtol or it >= maxit:
                break
        y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used])
        x = x + (np.array(V[:k_used]).T @ y) / d
    return x, it, float(np.linalg.norm(b - A @ x) / bn)


# ----------------------------------------------------------------------------
# a18-a19: field update and back-substitution
# ----------------------------------------------------------------------------
def field_update(n, free, x, fixed_mask, u_fixed):
    """SOLVER_VARIABLES_FIELD_UPDATE, solver_routines.F90:22814-23001: free DOFs take the
    solution (coupling coefficient 1, constant 0), fixed DOFs keep their BC values."""
    u = np.where(fixed_mask, np.asarray(u_fixed, dtype=np.float64), 0.0)
    u[free] = x
    return u


# ----------------------------------------------------------------------------
# Linear DOF constraints: coupled rows and columns in the solver mapping
# ----------------------------------------------------------------------------
def constraint_couplings(fixed_mask, constraints):
    """Solver DOF (owner) and coupling coefficient of every equations DOF.

    constraints: (c, m, a) triples, 0-based, meaning u_c = a u_m.  Follows
    BoundaryConditions_DofConstraintSet (boundary_condition_routines.F90:3204-3313:
    one DOF per constraint, the DOF depended on must be free), the check and
    coupling lists of BoundaryConditions_DofConstraintsCreateFinish (:3320-3472: a
    constrained DOF later fixed is an error; m's list is [m (1), c (a), ...]) and the
    row inclusion of SOLVER_MAPPING_CALCULATE (solver_mapping_routines.F90:422-520:
    only FREE rows are solver rows, in ascending order; a constrained row enters
    through the coupling list of its master).  A DOF constrained twice is rejected.
    Returns (owner int64 [n], -1 where fixed; coef float64 [n]).
    """
    fixed = np.asarray(fixed_mask, dtype=bool)
    n = len(fixed)
    master = np.full(n, -1, dtype=np.int64)
    coef = np.ones(n, dtype=np.float64)
    for c, m, a in constraints:
        c, m = int(c), int(m)
        if c == m:
            raise ValueError(f"DOF {c} is constrained to itself")
        if master[c] >= 0:
            raise ValueError(f"DOF {c} is constrained twice")
        if fixed[m] or master[m] >= 0:
            raise ValueError(f"DOF {m} is not free in the boundary conditions")
        if fixed[c]:
            raise ValueError(f"DOF {c} is part of a linear constraint but is fixed")
        master[c] = m
        coef[c] = float(a)
    if np.any(master[master[master >= 0]] >= 0):
        raise ValueError("a constraint depends on a constrained DOF")
    solver_dof = (~fixed) & (master < 0)
    rank = np.full(n, -1, dtype=np.int64)
    rank[solver_dof] = np.arange(int(solver_dof.sum()))
    owner = np.where(master >= 0, rank[np.maximum(master, 0)], rank)
    owner[fixed] = -1
    return owner, coef


def reduce_coupled(row_ptr, col, vals, owner, coef, u_fixed, rhs_extra, rhs_fixed_mask=None):
    """Solver system P^T K P and its RHS with coupled rows and columns.

    Matrix: SOLVER_MATRIX_EQUATIONS_MATRIX_ADD (solver_matrices_routines.F90:935-960)
    adds K_rc a_r a_c to solver entry (owner[r], owner[c]) for every equations entry,
    row by row; the pattern of SOLVER_MATRIX_STRUCTURE_CALCULATE (:1612-2431) is the
    sorted, unique set of those entries.  RHS: SOLVER_MATRICES_STATIC_ASSEMBLE
    (solver_routines.F90:14640-14850) adds a_r (rhs_extra_r - lift_r) to solver row
    owner[r], with the ZERO_TOLERANCE lifting of reduce_dirichlet over the fixed
    columns.  Returns (s_row_ptr, s_col, s_vals, b).
    """
    owner = np.asarray(owner, dtype=np.int64)
    coef = np.asarray(coef, dtype=np.float64)
    n = len(owner)
    n_s = int(owner.max()) + 1 if n and owner.max() >= 0 else 0
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    keep = (owner[rows] >= 0) & (owner[col] >= 0)
    key = owner[rows[keep]] * n_s + owner[col[keep]]
    contrib = (vals[keep] * coef[rows[keep]]) * coef[col[keep]]
    order = np.argsort(key, kind="stable")
    uk, first = np.unique(key[order], return_index=True)
    s_vals = np.add.reduceat(contrib[order], first) if len(first) else np.zeros(0)
    s_row = uk // max(n_s, 1)
    s_col = uk % max(n_s, 1)
    s_row_ptr = np.zeros(n_s + 1, dtype=np.int64)
    np.add.at(s_row_ptr, s_row + 1, 1)
    s_row_ptr = np.cumsum(s_row_ptr)
    fixed = owner < 0
    lift = (~fixed[rows]) & fixed[col]
    if rhs_fixed_mask is not None:
        lift &= ~np.asarray(rhs_fixed_mask, dtype=bool)[rows]
    ud = np.asarray(u_fixed, dtype=np.float64)[col[lift]]
    kv = vals[lift]
    ok = (np.abs(ud) >= ZERO_TOLERANCE) & (np.abs(kv) >= ZERO_TOLERANCE)
    r_eff = np.asarray(rhs_extra, dtype=np.float64).copy()
    np.add.at(r_eff, rows[lift][ok], -1.0 * kv[ok] * ud[ok])
    b = np.zeros(n_s)
    nf = ~fixed
    np.add.at(b, owner[nf], r_eff[nf] * coef[nf])
    return s_row_ptr, s_col, s_vals, b


def field_update_coupled(owner, coef, x, u_fixed):
    """SOLVER_VARIABLES_FIELD_UPDATE with coupling coefficients (solver_routines.F90:
    22814-23001, value = x * coefficient + 0 at :22877); fixed DOFs keep their values."""
    owner = np.asarray(owner, dtype=np.int64)
    u = np.where(owner < 0, np.asarray(u_fixed, dtype=np.float64), 0.0)
    nf = owner >= 0
    u[nf] = np.asarray(x, dtype=np.float64)[owner[nf]] * np.asarray(coef)[nf]
    return u


def backsubstitute(row_ptr, col, vals, u, source=None, rhs_fixed_mask=None, iron_compat=False):
    """EQUATIONS_SET_BACKSUBSTITUTE, CSR branch, equations_set_routines.F90:1784-1830.

    For rows whose RHS-variable DOF is free: R_i = sum_k K(i,col_k) * u[.] - s_i, where
    iron reads u at position (k - rowstart) of the row (:1806, restated when
    ``iron_compat``), while the correct form reads u[col_k].  Rows with a fixed RHS DOF
    get R_i = -s_i (the update at :1822 runs for every row).
    """
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    if iron_compat:
        k_in_row = np.arange(len(col)) - row_ptr[rows]
        uu = u[k_in_row]
    else:
        uu = u[col]
    R = np.zeros(n)
    np.add.at(R, rows, vals * uu)
    if rhs_fixed_mask is not None:
        R[rhs_fixed_mask] = 0.0
    if source is not None:
        R = R - source
    return R


# ----------------------------------------------------------------------------
# a20: Neumann point conditions integrated over boundary faces
# ----------------------------------------------------------------------------
def neumann_integrate(mesh: RegularMesh, dm: DofMap, point_values: dict, integrated_only=frozenset()):
    """BoundaryConditions_NeumannIntegrate, 3-D node-based branch,
    boundary_condition_routines.F90:2654-2996 (faces loop :2836-2920).

    For every NEUMANN_POINT DOF o (value q_o, node n, component c) and every
    boundary face that contains n, integrationMatrix(m, o) += sum_g phi_m phi_o
    w_g |J_area| over the face nodes m of component c (face basis = the 2-D
    Lagrange basis of the element's order, default Gauss scheme); then
    f = integrationMatrix . q.  Loops are written as the reference nests them.
    Single-basis meshes (the geometry basis equals the field basis)."""
    b3 = mesh.bases[0]
    fb = Basis(nxi=2, order=b3.order, ngauss_xi=b3.ngauss_xi)
    f = np.zeros(dm.num_dofs)
    E = mesh.num_elements
    ne = mesh.n_elem
    for o, q in point_values.items():
        c = int(np.searchsorted(np.array(dm.offsets[1:] + [dm.num_dofs]), o, side="right"))
        node_user = int(mesh.comp_nodes[dm.comp_mesh[c]][o - dm.offsets[c]])
        for e in range(E):
            ex = [(e // int(np.prod(ne[:k]))) % ne[k] for k in range(3)]
            en = mesh.elem_nodes[0][e]
            if node_user not in en:
                continue
            for k in range(3):
                for side in (0, 1):
                    if ex[k] != (0 if side == 0 else ne[k] - 1):
                        continue  # not a boundary face
                    want = 1 if side == 0 else b3.nodes_xic
                    face_local = [a for a in range(b3.number_of_nodes) if b3.node_index[a, k] == want]
                    face_nodes = en[face_local]
                    if node_user not in face_nodes:
                        continue
                    face_dofs = [dm.offsets[c] + dm.node_local[dm.comp_mesh[c]][nd] for nd in face_nodes]
                    if any(int(d) in integrated_only for d in face_dofs):
                        continue
                    ol = list(face_nodes).index(node_user)
                    xf = mesh.coords[face_nodes - 1]
                    tang = [j for j in range(3) if j != k]
                    for g in range(len(fb.gauss_weights)):
                        J = np.zeros((3, 2))
                        for a in range(len(face_nodes)):
                            for t in range(2):
                                J[:, t] += xf[a] * fb.gauss_basis_fns[a, FIRST_DERIV_MAP[t], g]
                        GL = J.T @ J
                        jac = math.sqrt(abs(GL[0, 0] * GL[1, 1] - GL[0, 1] * GL[1, 0]))
                        phio = fb.gauss_basis_fns[ol, NO_PART_DERIV, g]
                        for m in range(len(face_nodes)):
                            phim = fb.gauss_basis_fns[m, NO_PART_DERIV, g]
                            dof_m = dm.offsets[c] + dm.node_local[dm.comp_mesh[c]][face_nodes[m]]
                            f[dof_m] += q * phim * phio * fb.gauss_weights[g] * jac
                    del tang
    return f
