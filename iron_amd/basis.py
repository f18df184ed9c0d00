"""Lagrange tensor-product bases and Gauss-Legendre quadrature (host, FP64).

Tables are tiny (at most 27 nodes x 27 points) and built once per basis on the
host, then uploaded.  Conventions follow the reference so that element
matrices index the same way:
  * Gauss points and weights on [0,1] (basis_routines.F90:6804-6893);
  * element nodes xi1-fastest (basis_routines.F90:2551-2590);
  * Gauss points xi1-fastest, ng = i + (j-1 + (k-1) n2) n1 (basis_routines.F90:3813-3841).
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

LINEAR_LAGRANGE, QUADRATIC_LAGRANGE, CUBIC_LAGRANGE = 1, 2, 3
_ORDER = {LINEAR_LAGRANGE: 1, QUADRATIC_LAGRANGE: 2, CUBIC_LAGRANGE: 3}


def gauss_points_01(n: int):
    """Gauss-Legendre abscissae/weights mapped to [0,1] (1 <= n <= 4)."""
    if not 1 <= n <= 4:
        raise ValueError(f"The specified number of Gauss points of {n} is invalid or not implemented.")
    if n == 1:
        t, w = [0.0], [2.0]
    elif n == 2:
        a = 1.0 / math.sqrt(3.0)
        t, w = [-a, a], [1.0, 1.0]
    elif n == 3:
        a = math.sqrt(0.6)
        t, w = [-a, 0.0, a], [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]
    else:
        r = math.sqrt(6.0 / 5.0)
        a, b = math.sqrt((3.0 + 2.0 * r) / 7.0), math.sqrt((3.0 - 2.0 * r) / 7.0)
        wa, wb = (18.0 - math.sqrt(30.0)) / 36.0, (18.0 + math.sqrt(30.0)) / 36.0
        t, w = [-a, -b, b, a], [wa, wb, wb, wa]
    x = np.array([(ti + 1.0) / 2.0 for ti in t])
    return x, np.array(w) / 2.0


def _lagrange_1d(order: int, xi: np.ndarray):
    """Values and first derivatives of the 1-D Lagrange functions through
    equally spaced nodes on [0,1]; returns (phi[nodes, len(xi)], dphi[...])."""
    nodes = np.linspace(0.0, 1.0, order + 1)
    phi = np.ones((order + 1, len(xi)))
    dphi = np.zeros((order + 1, len(xi)))
    for a in range(order + 1):
        others = [b for b in range(order + 1) if b != a]
        den = np.prod([nodes[a] - nodes[b] for b in others])
        for b in others:
            phi[a] *= (xi - nodes[b])
        phi[a] /= den
        for c in others:
            term = np.ones(len(xi))
            for b in others:
                if b != c:
                    term = term * (xi - nodes[b])
            dphi[a] += term
        dphi[a] /= den
    return phi, dphi


class LagrangeBasis:
    """Tensor-product Lagrange basis with its default Gauss scheme."""

    def __init__(self, num_xi: int, interpolation: int, gauss_xi: int | None = None):
        if interpolation not in _ORDER:
            raise ValueError(f"interpolation {interpolation} is not a Lagrange interpolation")
        self.num_xi = num_xi
        self.interpolation = interpolation
        self.order = _ORDER[interpolation]
        self.nodes_xi = self.order + 1
        self.gauss_xi = gauss_xi if gauss_xi is not None else self.order + 1
        self.num_nodes = self.nodes_xi ** num_xi
        self.num_gauss = self.gauss_xi ** num_xi
        self._tabulate()

    def _tabulate(self):
        d, n1, g1 = self.num_xi, self.nodes_xi, self.gauss_xi
        x1, w1 = gauss_points_01(g1)
        phi1, dphi1 = _lagrange_1d(self.order, x1)  # [n1, g1]
        # multi-indices, first index fastest
        node_idx = np.array(np.unravel_index(np.arange(n1 ** d), (n1,) * d, order="F")).T  # [nn, d]
        gp_idx = np.array(np.unravel_index(np.arange(g1 ** d), (g1,) * d, order="F")).T    # [ng, d]
        nn, ng = len(node_idx), len(gp_idx)
        N = np.ones((ng, nn))
        dN = np.ones((ng, nn, d))
        W = np.ones(ng)
        for k in range(d):
            a = node_idx[:, k][None, :]
            g = gp_idx[:, k][:, None]
            N *= phi1[a, g]
            W *= w1[gp_idx[:, k]]
            for j in range(d):
                dN[:, :, j] *= (dphi1 if j == k else phi1)[a, g]
        self.node_index = node_idx
        self.gauss_positions = x1[gp_idx]
        self.N = N            # [ng, nn]
        self.dN = dN          # [ng, nn, d]
        self.weights = W      # [ng]

    def corner_nodes(self):
        """Local indices of the element vertices (first/last node per direction)."""
        last = self.nodes_xi - 1
        mask = np.all((self.node_index == 0) | (self.node_index == last), axis=1)
        return np.nonzero(mask)[0]


@lru_cache(maxsize=None)
def basis(num_xi: int, interpolation: int, gauss_xi: int | None = None) -> LagrangeBasis:
    return LagrangeBasis(num_xi, interpolation, gauss_xi)


def face_quadrature(b: LagrangeBasis, face_xi: int, face_value: int):
    """Quadrature on an element face xi_{face_xi} = face_value (0 or 1) using the
    basis' Gauss points in the remaining directions.  Returns (N[ngf, nn],
    dN[ngf, nn, d], wf[ngf])."""
    d = b.num_xi
    x1, w1 = gauss_points_01(b.gauss_xi)
    others = [k for k in range(d) if k != face_xi]
    pts = np.array(np.unravel_index(np.arange(b.gauss_xi ** (d - 1)), (b.gauss_xi,) * (d - 1), order="F")).T
    ngf = len(pts)
    xi = np.zeros((ngf, d))
    wf = np.ones(ngf)
    for c, k in enumerate(others):
        xi[:, k] = x1[pts[:, c]]
        wf *= w1[pts[:, c]]
    xi[:, face_xi] = float(face_value)
    nodes1 = np.linspace(0.0, 1.0, b.nodes_xi)
    N = np.ones((ngf, b.num_nodes))
    dN = np.ones((ngf, b.num_nodes, d))
    for k in range(d):
        phi, dphi = _lagrange_1d(b.order, xi[:, k])
        a = b.node_index[:, k][None, :]
        g = np.arange(ngf)[:, None]
        N *= phi[a, g]
        for j in range(d):
            dN[:, :, j] *= (dphi if j == k else phi)[a, g]
    del nodes1
    return N, dN, wf
