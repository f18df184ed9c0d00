"""Regular generated meshes and separated DOF numbering (host, vectorised).

Numbering follows the reference exactly, because the CSR pattern and the DOF
map must come out identical to iron's:
  * elements xi1-fastest; element node user numbers on the first basis' node
    lattice (generated_mesh_routines.F90:1568-1612);
  * a coarser second basis (e.g. trilinear pressure under triquadratic
    displacement) takes the first basis' corner user numbers
    (GeneratedMesh_RegularComponentNodesToUserNumbers, :4941-5357);
  * each mesh component's nodes are its sorted unique user numbers, local
    number = rank (mesh_routines.F90:9159-9279, one domain :4603-4620);
  * coordinates origin + i * extent/(n-1) per direction
    (GeneratedMesh_RegularGeometricParametersCalculate, :3217-3384);
  * field DOFs in SEPARATED order: component-major, node-minor
    (FIELD_MAPPINGS_CALCULATE, field_routines.F90:10056).
"""
from __future__ import annotations

import numpy as np

from .basis import LagrangeBasis


def _lattice(dims):
    """[prod(dims), len(dims)] multi-indices, first index fastest (Fortran order)."""
    dims = [int(d) for d in dims]
    total = int(np.prod(dims))
    idx = np.arange(total, dtype=np.int64)
    cols, inner = [], 1
    for d in dims:
        cols.append((idx // inner) % d)
        inner *= d
    return np.stack(cols, axis=1)


def _outer_sum(terms):
    """Sum over a Fortran-ordered grid: terms[k] are 1-D arrays along direction k
    (first fastest); returns the flattened [prod(len)] array."""
    out = terms[-1]
    for t in terms[-2::-1]:
        out = np.add.outer(out, t)
    return out.reshape(-1)


class RegularMesh:
    def __init__(self, n_elem, bases, extent, origin=None):
        self.n_elem = tuple(int(n) for n in n_elem)
        self.dim = len(self.n_elem)
        self.bases = list(bases)
        self.extent = tuple(float(e) for e in extent)
        self.origin = tuple(float(o) for o in (origin if origin is not None else (0.0,) * self.dim))
        if any(b.num_xi != self.dim for b in self.bases):
            raise ValueError("basis dimension does not match the mesh")
        self.num_elements = int(np.prod(self.n_elem))
        self._build()

    def _build(self):
        d, ne = self.dim, self.n_elem
        b0 = self.bases[0]
        p0 = b0.nodes_xi - 1
        tot0 = [p0 * n + 1 for n in ne]
        strides0 = np.cumprod([1] + tot0[:-1])
        self.user_nodes = []     # [b] -> [E, nn] user numbers (1-based; the first basis' only with a second basis)
        self.comp_user = []      # [b] -> sorted unique user numbers of the mesh component
        self.elem_nodes = []     # [b] -> [E, nn] local node numbers (0-based)
        for bi, b in enumerate(self.bases):
            if bi == 0:
                # first node - 1 of every element: sum over directions of i_k p0 stride_k
                base = _outer_sum([np.arange(ne[k], dtype=np.int32) * np.int32(p0 * strides0[k]) for k in range(d)])
                offs = (b.node_index * strides0[None, :]).sum(axis=1).astype(np.int32)  # [nn]
                # every lattice node of the first basis belongs to an element, so its
                # sorted unique user numbers are 1..prod(tot0) and local = user - 1
                local = base[:, None] + offs[None, :]
                self.user_nodes.append(local.astype(np.int64) + 1 if len(self.bases) > 1 else None)
                self.comp_user.append(np.arange(1, int(np.prod(tot0)) + 1, dtype=np.int64))
                self.elem_nodes.append(local)
                continue
            elif b.nodes_xi == b0.nodes_xi:
                user = self.user_nodes[0].copy()
            elif b.nodes_xi == 2:
                corner = np.where(b.node_index == 0, 0, b0.nodes_xi - 1)  # [nn, d] first-basis local idx
                a0 = (corner * np.cumprod([1] + [b0.nodes_xi] * (d - 1))[None, :]).sum(axis=1)
                user = self.user_nodes[0][:, a0]
            else:
                raise NotImplementedError("secondary bases must match the first or be linear")
            uniq = np.unique(user)
            self.user_nodes.append(user.astype(np.int64))
            self.comp_user.append(uniq)
            self.elem_nodes.append(np.searchsorted(uniq, user).astype(np.int32))
        # geometric coordinates on mesh component 1
        delta = [self.extent[k] / float(tot0[k] - 1) for k in range(d)]
        pos = _lattice(tot0)  # [N, d]: component-1 nodes are the whole lattice, in order
        coords = np.empty((len(self.comp_user[0]), d))
        for c in range(d):
            # origin + i * delta, evaluated once per lattice line and then broadcast
            line = self.origin[c] + np.arange(tot0[c], dtype=np.int64).astype(np.float64) * delta[c]
            coords[:, c] = line[pos[:, c]]
        self.coords = coords
        self.node_grid_pos = pos

    def num_nodes(self, comp: int = 0) -> int:
        return len(self.comp_user[comp])

    def local_node(self, user_number, comp: int = 0):
        u = np.asarray(user_number)
        loc = np.searchsorted(self.comp_user[comp], u)
        if np.any(loc >= len(self.comp_user[comp])) or np.any(self.comp_user[comp][np.minimum(loc, len(self.comp_user[comp]) - 1)] != u):
            raise ValueError("node user number not in mesh component")
        return loc

    def node_coords(self, comp: int = 0) -> np.ndarray:
        """Coordinates of the nodes of mesh component ``comp`` (its nodes are a
        subset of component 1's user numbers)."""
        if comp == 0:
            return self.coords
        return self.coords[np.searchsorted(self.comp_user[0], self.comp_user[comp])]


class DofMap:
    """Separated DOF numbering of a field variable whose component c lives on
    mesh component comp_mesh[c] (node-based Lagrange: one DOF per node)."""

    def __init__(self, mesh: RegularMesh, comp_mesh):
        self.mesh = mesh
        self.comp_mesh = list(comp_mesh)
        self.offsets = []
        off = 0
        for mc in self.comp_mesh:
            self.offsets.append(off)
            off += mesh.num_nodes(mc)
        self.num_dofs = off

    def dof(self, c: int, local_node):
        return self.offsets[c] + np.asarray(local_node)

    def element_dofs(self) -> np.ndarray:
        """[E, n_edof] component-major, then element node (int32)."""
        E = self.mesh.num_elements
        widths = [self.mesh.elem_nodes[mc].shape[1] for mc in self.comp_mesh]
        out = np.empty((E, sum(widths)), dtype=np.int32)
        k = 0
        for c, mc in enumerate(self.comp_mesh):
            np.add(self.mesh.elem_nodes[mc], np.int32(self.offsets[c]), out=out[:, k:k + widths[c]])
            k += widths[c]
        return out

    def component_of(self, dof: np.ndarray) -> np.ndarray:
        return np.searchsorted(np.array(self.offsets[1:] + [self.num_dofs]), dof, side="right")


def structured_colours(mesh: RegularMesh):
    """Element colours of a regular grid: parity of the element index in each
    direction (2^dim colours; no two elements of a colour share a node)."""
    return _outer_sum([(np.arange(n, dtype=np.int32) % 2) << k for k, n in enumerate(mesh.n_elem)]).astype(np.int32)
