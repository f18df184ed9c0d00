"""The oracle's solver mapping with linear DOF constraints (u_c = a u_m).

What pins it, independently of the restatement's own formula:
  * with no constraints it reproduces the Dirichlet reduction entry for entry;
  * its solution satisfies the Lagrange-multiplier (KKT) form of the same
    constrained energy minimum, solved densely with the constraints as rows of B;
  * its matrix equals the dense product P^T K P with P built column by column;
  * tying the DOFs of opposite faces of the unit cube (coefficient 1) keeps the
    closed-form solution u = x of the Laplace problem with u = 0 / 1 on x = 0 / 1.
"""
import numpy as np
import pytest

from oracle import fe_oracle as fo


def _laplace(ne, order=1, seed=0, source=None, uniform=False):
    rng = np.random.default_rng(seed)
    ob = fo.Basis(nxi=3, order=order, ngauss_xi=order + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0])
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    kappa = np.ones(om.num_elements) if uniform else np.exp(rng.standard_normal(om.num_elements))
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    src = np.zeros(od.num_dofs) if source is None else \
        fo.assemble_vector(od.element_dofs(), fo.poisson_source_vectors(xe, ob, source), od.num_dofs)
    return om.coords, rp, col, vals, src


def _face_ties(X, a_of):
    """(c, m, a): each node on y = 1 tied to the node on y = 0 with the same x, z,
    skipping the Dirichlet faces x = 0 and x = 1."""
    key = {(round(x, 9), round(z, 9)): i for i, (x, y, z) in enumerate(X) if abs(y) < 1e-12}
    out = []
    for i, (x, y, z) in enumerate(X):
        if abs(y - 1.0) < 1e-12 and 1e-12 < x < 1 - 1e-12:
            m = key[(round(x, 9), round(z, 9))]
            out.append((i, m, a_of(i)))
    return out


def _dense(rp, col, vals, n):
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    A[rows, col] = vals
    return A


def test_no_constraints_reproduce_the_dirichlet_reduction():
    X, rp, col, vals, src = _laplace((3, 2, 2), order=2)
    fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1) < 1e-12)
    ufix = np.where(np.abs(X[:, 0] - 1) < 1e-12, 1.0, 0.0)
    owner, coef = fo.constraint_couplings(fixed, [])
    free, srp, scol, sv, b = fo.reduce_dirichlet(rp, col, vals, fixed, ufix, src)
    crp, ccol, cv, cb = fo.reduce_coupled(rp, col, vals, owner, coef, ufix, src)
    assert np.array_equal(crp, srp) and np.array_equal(ccol, scol)
    assert np.array_equal(cv, sv) and np.array_equal(cb, b)
    assert np.array_equal(owner[free], np.arange(len(free)))


@pytest.mark.parametrize("order", [1, 2])
def test_coupled_matrix_is_ptkp_and_solution_solves_the_kkt_system(order):
    X, rp, col, vals, src = _laplace((3, 3, 2), order=order, seed=3, source=1.0)
    n = len(X)
    fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1) < 1e-12)
    ufix = np.where(np.abs(X[:, 0]) < 1e-12, 0.3, 1.0)
    rng = np.random.default_rng(7)
    coefs = {}
    ties = _face_ties(X, lambda i: coefs.setdefault(i, float(rng.uniform(0.5, 1.5))))
    assert len(ties) > 4
    owner, coef = fo.constraint_couplings(fixed, ties)
    srp, scol, sv, b = fo.reduce_coupled(rp, col, vals, owner, coef, ufix, src)
    n_s = len(srp) - 1
    assert n_s == n - int(fixed.sum()) - len(ties)
    # rows sorted and unique
    for r in range(n_s):
        cr = scol[srp[r]:srp[r + 1]]
        assert np.all(np.diff(cr) > 0)
    # P^T K P with P[g, owner g] = coef g, and the pattern of P^T |K| P
    K = _dense(rp, col, vals, n)
    P = np.zeros((n, n_s))
    nf = owner >= 0
    P[np.nonzero(nf)[0], owner[nf]] = coef[nf]
    S = _dense(srp, scol, sv, n_s)
    np.testing.assert_allclose(S, P.T @ K @ P, rtol=0, atol=1e-13 * np.abs(K).max())
    pattern = (np.abs(P).T @ (K != 0) @ np.abs(P)) != 0
    assert np.array_equal(S != 0, pattern)
    x = np.linalg.solve(S, b)
    u = fo.field_update_coupled(owner, coef, x, ufix)
    # KKT: K u - B^T lam = f, B u = g (fixed values and u_c - a u_m = 0)
    rows_B, g = [], []
    for d in np.nonzero(fixed)[0]:
        e = np.zeros(n); e[d] = 1.0; rows_B.append(e); g.append(ufix[d])
    for c, m, a in ties:
        e = np.zeros(n); e[c] = 1.0; e[m] = -a; rows_B.append(e); g.append(0.0)
    B = np.array(rows_B)
    kkt = np.block([[K, B.T], [B, np.zeros((len(B), len(B)))]])
    sol = np.linalg.solve(kkt, np.concatenate([src, g]))
    np.testing.assert_allclose(u, sol[:n], rtol=1e-10, atol=1e-12)
    for c, m, a in ties:
        assert u[c] == x[owner[m]] * a


def test_periodic_faces_keep_the_linear_closed_form():
    X, rp, col, vals, src = _laplace((4, 4, 4), order=1, uniform=True)
    fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1) < 1e-12)
    ufix = X[:, 0].copy()
    ties = _face_ties(X, lambda i: 1.0)
    owner, coef = fo.constraint_couplings(fixed, ties)
    srp, scol, sv, b = fo.reduce_coupled(rp, col, vals, owner, coef, ufix, src)
    x = np.linalg.solve(_dense(srp, scol, sv, len(srp) - 1), b)
    u = fo.field_update_coupled(owner, coef, x, ufix)
    np.testing.assert_allclose(u, X[:, 0], rtol=0, atol=1e-12)


def test_constraint_rules_of_the_reference():
    fixed = np.array([True, False, False, False, False])
    with pytest.raises(ValueError):  # the DOF depended on must be free
        fo.constraint_couplings(fixed, [(1, 0, 1.0)])
    with pytest.raises(ValueError):  # a constrained DOF that is also fixed
        fo.constraint_couplings(fixed, [(0, 1, 1.0)])
    with pytest.raises(ValueError):  # constrained twice
        fo.constraint_couplings(fixed, [(2, 1, 1.0), (2, 3, 1.0)])
    with pytest.raises(ValueError):  # depending on a constrained DOF
        fo.constraint_couplings(fixed, [(2, 1, 1.0), (3, 2, 1.0)])
    owner, coef = fo.constraint_couplings(fixed, [(4, 1, -2.0), (2, 1, 0.5)])
    assert owner.tolist() == [-1, 0, 0, 1, 0] and coef.tolist() == [1.0, 1.0, 0.5, 1.0, -2.0]
