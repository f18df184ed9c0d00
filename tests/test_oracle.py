"""Pins the FP64 oracle (oracle/fe_oracle.py) against the reference's golden
output and against closed forms, before it is trusted as the GPU's checker."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import fe_oracle as fo

GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ a4
@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_gauss_exact_integration(n):
    x, w = fo.gauss_legendre(n)
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-15
    # one degree more is not exact (pins the number of points)
    assert abs(np.sum(w * x ** (2 * n)) - 1.0 / (2 * n + 1)) > 1e-6


@pytest.mark.parametrize("nxi,order", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
def test_basis_partition_of_unity_and_nodal(nxi, order):
    b = fo.Basis(nxi=nxi, order=order, ngauss_xi=order + 1)
    tab = b.gauss_basis_fns
    assert np.allclose(tab[:, 1, :].sum(axis=0), 1.0, atol=1e-14)
    for nu in fo.FIRST_DERIV_MAP[:nxi]:
        assert np.allclose(tab[:, nu, :].sum(axis=0), 0.0, atol=1e-13)
    # Kronecker property at the nodes
    for a in range(b.number_of_nodes):
        xi = (b.node_index[a] - 1) / order
        vals = [b.evaluate(c, 1, xi) for c in range(b.number_of_nodes)]
        assert np.allclose(vals, np.eye(b.number_of_nodes)[a], atol=1e-14)


# --------------------------------------------------------- a1-a3, a14
def _dense_pattern(edofs, n):
    M = np.zeros((n, n), dtype=bool)
    for row in edofs:
        M[np.ix_(row, row)] = True
    return M


@pytest.mark.parametrize("ne,order,ncomp", [((3, 2, 2), 1, 1), ((2, 2, 1), 2, 1), ((2, 1, 2), 1, 3), ((3, 2), 2, 1)])
def test_csr_pattern_matches_element_union(ne, order, ncomp):
    b = fo.Basis(nxi=len(ne), order=order, ngauss_xi=order + 1)
    m = fo.regular_mesh(ne, [b], (1.0,) * len(ne))
    dm = fo.DofMap(m, [0] * ncomp)
    rp, col = fo.csr_pattern(dm)
    M = _dense_pattern(dm.element_dofs(), dm.num_dofs)
    assert rp[-1] == M.sum()
    for i in range(dm.num_dofs):
        assert np.array_equal(col[rp[i]:rp[i + 1]], np.nonzero(M[i])[0])
    # 1-D factor products: Q1 3n+1 per direction (n+1 nodes), Q2 8n+1
    per = [(3 * n + 1) if order == 1 else (8 * n + 1) for n in ne]
    assert rp[-1] == int(np.prod(per)) * ncomp * ncomp


def test_mesh_numbering_hand_enumerated():
    # 2x1x1 trilinear: nodes x-fastest on a 3x2x2 lattice
    b = fo.Basis(nxi=3, order=1, ngauss_xi=2)
    m = fo.regular_mesh((2, 1, 1), [b], (2.0, 1.0, 1.0))
    assert m.elem_nodes[0].tolist() == [[1, 2, 4, 5, 7, 8, 10, 11], [2, 3, 5, 6, 8, 9, 11, 12]]
    assert np.allclose(m.coords[1], [1.0, 0.0, 0.0]) and np.allclose(m.coords[11], [2.0, 1.0, 1.0])
    # mixed triquadratic / trilinear: linear nodes take the quadratic corner numbers
    bq = fo.Basis(nxi=3, order=2, ngauss_xi=3)
    mm = fo.regular_mesh((1, 1, 1), [bq, b], (1.0, 1.0, 1.0))
    assert mm.elem_nodes[1].tolist() == [[1, 3, 7, 9, 19, 21, 25, 27]]
    dm = fo.DofMap(mm, [0, 0, 0, 1])
    assert dm.num_dofs == 3 * 27 + 8
    assert dm.dof(3, [1, 27]).tolist() == [81, 88]


def test_exnode_coordinates_match_generated_mesh():
    # Cantilever linear_3x2x2 golden: 3x2x2 elements on a 60x40x40 box; node 2 at x = 20
    b = fo.Basis(nxi=3, order=1, ngauss_xi=2)
    m = fo.regular_mesh((3, 2, 2), [b], (60.0, 40.0, 40.0))
    assert np.allclose(m.coords[1], [20.0, 0.0, 0.0])
    assert np.allclose(m.coords[-1], [60.0, 40.0, 40.0])


# ---------------------------------------------------------- a6-a9
def _box_elements(ne, order, extent):
    b = fo.Basis(nxi=len(ne), order=order, ngauss_xi=order + 1)
    m = fo.regular_mesh(ne, [b], extent)
    xe = m.coords[m.elem_nodes[0] - 1]
    return b, m, xe


def test_q1_laplace_unit_cube_closed_form():
    b, m, xe = _box_elements((1, 1, 1), 1, (1.0, 1.0, 1.0))
    K = fo.laplace_element_matrices(xe, b)[0]
    # trilinear unit cube: 4/12 diagonal, 0 along edges, -1/12 face and body diagonals
    for a in range(8):
        for c in range(8):
            dist = int(np.sum(b.node_index[a] != b.node_index[c]))
            expect = {0: 4 / 12, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[dist]
            assert abs(K[a, c] - expect) < 1e-15


def test_q1_laplace_2d_closed_form_and_scaling():
    b, m, xe = _box_elements((1, 1), 1, (1.0, 1.0))
    K = fo.laplace_element_matrices(xe, b)[0]
    assert abs(K[0, 0] - 2 / 3) < 1e-15 and abs(K[0, 1] + 1 / 6) < 1e-15 and abs(K[0, 3] + 1 / 3) < 1e-15
    # 3-D scaling K(h) = h K(1); generalised Laplace with kappa*I = kappa * standard
    b3, _, xe1 = _box_elements((1, 1, 1), 2, (1.0, 1.0, 1.0))
    _, _, xe2 = _box_elements((1, 1, 1), 2, (0.25, 0.25, 0.25))
    K1 = fo.laplace_element_matrices(xe1, b3)[0]
    K2 = fo.laplace_element_matrices(xe2, b3)[0]
    assert np.allclose(K2, 0.25 * K1, rtol=1e-13, atol=1e-15)
    K3 = fo.laplace_element_matrices(xe1, b3, kappa=[2.5])[0]
    assert np.allclose(K3, 2.5 * K1, rtol=1e-14)
    assert np.allclose(K1.sum(axis=1), 0.0, atol=1e-13) and np.allclose(K1, K1.T, atol=1e-15)


def test_poisson_source_integrates_volume():
    b, m, xe = _box_elements((2, 1, 1), 2, (2.0, 0.5, 0.75))
    s = fo.poisson_source_vectors(xe, b, 3.0)
    assert abs(s.sum() - 3.0 * 2.0 * 0.5 * 0.75) < 1e-13


def test_isotropic_tensor_matches_lame():
    E, nu = 7.0, 0.3
    C = fo.linear_elasticity_tensor([E], [E], [E], [nu], [nu], [nu])[0]
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    assert np.allclose(np.diag(C), [lam + 2 * mu] * 3 + [mu] * 3, rtol=1e-14)
    assert np.allclose([C[0, 1], C[0, 2], C[1, 2]], lam, rtol=1e-14)


def test_elasticity_rigid_body_modes_and_patch():
    b, m, xe = _box_elements((1, 1, 1), 1, (2.0, 1.0, 0.5))
    C = fo.linear_elasticity_tensor([10.0], [10.0], [10.0], [0.3], [0.3], [0.3])
    K = fo.linear_elasticity_element_matrices(xe, b, C)[0]
    X = xe[0]
    modes = []
    for k in range(3):
        t = np.zeros((8, 3)); t[:, k] = 1.0
        modes.append(t)
    for (i, j) in [(0, 1), (0, 2), (1, 2)]:
        r = np.zeros((8, 3)); r[:, i] = -X[:, j]; r[:, j] = X[:, i]
        modes.append(r)
    for mode in modes:
        v = mode.T.ravel()  # component-blocked
        assert np.linalg.norm(K @ v) < 1e-12 * np.linalg.norm(K)
    assert np.allclose(K, K.T, atol=1e-12)
    assert np.linalg.matrix_rank(K, tol=1e-10 * np.linalg.norm(K)) == 24 - 6


# --------------------------------------------- a13, a16-a19 + golden file
def _analytic_laplace_problem():
    data = json.loads((GOLDEN / "analytic_laplace_linear_lagrange.json").read_text())
    o = -data["setup"]["origin_pi"] / 2.0
    b = fo.Basis(nxi=2, order=1, ngauss_xi=2)
    m = fo.regular_mesh((2, 2), [b], (2.0, 2.0), (o, o))
    dm = fo.DofMap(m, [0])
    rp, col = fo.csr_pattern(dm)
    xe = m.coords[m.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, dm.element_dofs(), fo.laplace_element_matrices(xe, b), dm.num_dofs)
    ua = np.cos(m.coords[:, 0]) * np.cosh(m.coords[:, 1])
    fixed = np.ones(9, bool)
    fixed[4] = False
    return data, rp, col, vals, ua, fixed


def _sig5(v):
    return float(f"{v:.4e}")


def test_golden_analytic_laplace_solution_and_compat_reactions():
    data, rp, col, vals, ua, fixed = _analytic_laplace_problem()
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, ua, np.zeros(9))
    x, it, rr = fo.cg_jacobi(srp, scol, sv, b, rtol=1e-14)
    u = fo.field_update(9, free, x, fixed, ua)
    phi = data["nodal"]["Phi"]
    for node, (num, ana) in phi.items():
        assert _sig5(u[int(node) - 1]) == pytest.approx(num, rel=1e-4, abs=1e-15)
        assert _sig5(ua[int(node) - 1]) == pytest.approx(ana, rel=1e-4, abs=1e-15)
    R_compat = fo.backsubstitute(rp, col, vals, u, iron_compat=True)
    R_true = fo.backsubstitute(rp, col, vals, u, iron_compat=False)
    flux = data["nodal"]["del Phi/del n"]
    mismatches = 0
    for node, (num, _) in flux.items():
        i = int(node) - 1
        if abs(num) < 1e-10:
            assert abs(R_compat[i]) < 1e-12
            continue
        assert R_compat[i] == pytest.approx(num, rel=1e-4)
        mismatches += abs(R_true[i] - num) > 1e-3 * abs(num)
    # the golden values come from the reference's in-row index (equations_set_routines.F90:1806)
    assert mismatches >= 4


def test_unit_cube_laplace_closed_form_u_equals_x():
    n = 4
    b = fo.Basis(nxi=3, order=1, ngauss_xi=2)
    m = fo.regular_mesh((n, n, n), [b], (1.0, 1.0, 1.0))
    dm = fo.DofMap(m, [0])
    rp, col = fo.csr_pattern(dm)
    xe = m.coords[m.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, dm.element_dofs(), fo.laplace_element_matrices(xe, b), dm.num_dofs)
    x0 = m.coords[:, 0]
    fixed = (np.abs(x0) < 1e-12) | (np.abs(x0 - 1.0) < 1e-12)
    ufix = np.where(np.abs(x0 - 1.0) < 1e-12, 1.0, 0.0)
    free, srp, scol, sv, bb = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, np.zeros(dm.num_dofs))
    x, it, rr = fo.cg_jacobi(srp, scol, sv, bb, rtol=1e-12)
    u = fo.field_update(dm.num_dofs, free, x, fixed, ufix)
    assert np.max(np.abs(u - x0)) < 1e-10
    # reactions (correct form): outward flux -1 on x=0 face nodes sum, +1 on x=1
    R = fo.backsubstitute(rp, col, vals, u)
    assert abs(R[np.abs(x0) < 1e-12].sum() + 1.0) < 1e-9
    assert abs(R[np.abs(x0 - 1) < 1e-12].sum() - 1.0) < 1e-9
    assert np.max(np.abs(R[~fixed])) < 1e-9


def test_krylov_against_dense_solve():
    rng = np.random.default_rng(1)
    b = fo.Basis(nxi=3, order=2, ngauss_xi=3)
    m = fo.regular_mesh((3, 2, 2), [b], (1.0, 1.0, 1.0))
    dm = fo.DofMap(m, [0])
    rp, col = fo.csr_pattern(dm)
    xe = m.coords[m.elem_nodes[0] - 1]
    kap = np.exp(rng.standard_normal(m.num_elements))
    vals = fo.assemble_csr(rp, col, dm.element_dofs(), fo.laplace_element_matrices(xe, b, kap), dm.num_dofs)
    src = fo.assemble_vector(dm.element_dofs(), fo.poisson_source_vectors(xe, b, 1.0), dm.num_dofs)
    X = m.coords
    fixed = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1.0) < 1e-12), axis=1)
    free, srp, scol, sv, bb = fo.reduce_dirichlet(rp, col, vals, fixed, np.zeros(dm.num_dofs), src)
    A = np.zeros((len(free), len(free)))
    for i in range(len(free)):
        A[i, scol[srp[i]:srp[i + 1]]] = sv[srp[i]:srp[i + 1]]
    xd = np.linalg.solve(A, bb)
    xc, _, rc = fo.cg_jacobi(srp, scol, sv, bb, rtol=1e-12)
    xg, _, rg = fo.gmres_jacobi(srp, scol, sv, bb, rtol=1e-12)
    assert rc <= 1e-12 and rg <= 1e-12
    assert np.allclose(xc, xd, rtol=1e-9, atol=1e-13) and np.allclose(xg, xd, rtol=1e-9, atol=1e-13)


def test_storage_location_find_bisection():
    rp = np.array([0, 30])
    col = np.arange(0, 60, 2)
    for j in range(60):
        p = fo.storage_location_find(rp, col, 0, j)
        assert p == (j // 2 if j % 2 == 0 else -1)


def test_dirichlet_lifting_zero_tolerance():
    # a fixed value below ZERO_TOLERANCE is not lifted (solver_routines.F90:14718)
    rp = np.array([0, 2, 4])
    col = np.array([0, 1, 0, 1])
    vals = np.array([2.0, -1.0, -1.0, 2.0])
    fixed = np.array([False, True])
    _, _, _, _, b = fo.reduce_dirichlet(rp, col, vals, fixed, np.array([0.0, 1e-16]), np.zeros(2))
    assert b[0] == 0.0
    _, _, _, _, b = fo.reduce_dirichlet(rp, col, vals, fixed, np.array([0.0, 0.5]), np.zeros(2))
    assert b[0] == 0.5
