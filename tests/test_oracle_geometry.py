"""The oracle's element matrices on geometry that is not an axis-aligned box.

On a box the Jacobian is diagonal, so a transposed metric index, a wrong
inverse or a misplaced material block can hide.  These checks use facts of the
mathematics, not the oracle's own formulas:
  * patch test: for a linear field u = a.x, K u vanishes on every row whose
    node is not on the boundary (the exact solution of the homogeneous problem
    is reproduced), on any isoparametric mesh whose integrals the Gauss rule
    evaluates exactly;
  * energy: u^T K u = integral of grad(u).kappa.grad(u) = kappa |a|^2 V;
  * linear elasticity: v^T K v = V e^T C e for v = G x (constant strain e in
    iron's Voigt order), and rigid motions in the null space.
Moving only interior nodes of a trilinear mesh keeps the domain (and its volume)
and keeps the 2x2x2 Gauss rule exact for these integrands (degree <= 3 per xi);
quadratic meshes are mapped by a global affine map, which keeps them exact.
"""
import numpy as np
import pytest

from oracle import fe_oracle as fo


def _mesh(ne, order, extent, ncomp=1):
    ob = fo.Basis(nxi=len(ne), order=order, ngauss_xi=order + 1)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0] * ncomp)
    return ob, om, od


def _boundary(X, extent):
    hi = np.asarray(extent)[None, :]
    return np.any((np.abs(X) < 1e-12) | (np.abs(X - hi) < 1e-12), axis=1)


def _distort_interior(X, extent, ne, rng, amp=0.2):
    """Random moves of the interior nodes by up to amp x element size per axis."""
    h = np.asarray(extent) / np.asarray(ne)
    Y = X.copy()
    inner = ~_boundary(X, extent)
    Y[inner] += amp * h[None, :] * rng.uniform(-1.0, 1.0, size=(int(inner.sum()), X.shape[1]))
    return Y


def _assemble_scalar(ob, om, od, coords, kappa):
    rp, col = fo.csr_pattern(od)
    xe = coords[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.laplace_element_matrices(xe, ob, kappa), od.num_dofs)
    return fo._csr_scipy(rp, col, vals)


def test_laplace_patch_test_and_energy_on_distorted_trilinear_mesh():
    rng = np.random.default_rng(21)
    ne, extent = (4, 3, 3), (1.0, 1.5, 2.0)
    ob, om, od = _mesh(ne, 1, extent)
    X = _distort_interior(om.coords, extent, ne, rng)
    assert np.abs(X - om.coords).max() > 0.05
    kappa = 2.5
    K = _assemble_scalar(ob, om, od, X, np.full(om.num_elements, kappa))
    a = np.array([0.7, -1.3, 0.4])
    u = X @ a
    r = K @ u
    inner = ~_boundary(X, extent)
    assert np.abs(r[inner]).max() < 1e-12 * np.abs(r).max()
    assert np.abs(r[~inner]).max() > 1e-3           # the boundary carries the flux
    V = np.prod(extent)
    assert u @ (K @ u) == pytest.approx(kappa * (a @ a) * V, rel=1e-12)
    # constants are in the null space, the matrix is symmetric
    assert np.abs(K @ np.ones(len(u))).max() < 1e-12 * kappa
    assert abs(K - K.T).max() < 1e-14 * abs(K).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_laplace_quadratic_on_affine_mesh(dim):
    """A general (non-symmetric) affine map of a triquadratic / biquadratic mesh."""
    ne = (3, 2, 2)[:dim]
    extent = (1.0, 2.0, 1.0)[:dim]
    ob, om, od = _mesh(ne, 2, extent)
    A = np.array([[1.0, 0.35, -0.2], [0.1, 0.8, 0.25], [-0.15, 0.3, 1.2]])[:dim, :dim]
    X = om.coords @ A.T
    K = _assemble_scalar(ob, om, od, X, np.ones(om.num_elements))
    a = np.array([0.3, -0.9, 1.1])[:dim]
    u = X @ a
    inner = ~_boundary(om.coords, extent)
    r = K @ u
    assert np.abs(r[inner]).max() < 1e-12 * np.abs(r).max()
    assert u @ r == pytest.approx((a @ a) * abs(np.linalg.det(A)) * np.prod(extent), rel=1e-12)


def _voigt_strain(G):
    """iron's Voigt order of the element tensor: (e11, e22, e33, g13, g23, g12); the
    shear pairs follow the diagonal blocks [C11,C66,C44 | C66,C22,C55 | C44,C55,C33]
    of linear_elasticity_routines.F90:1255-1275."""
    e = 0.5 * (G + G.T)
    return np.array([e[0, 0], e[1, 1], e[2, 2], 2 * e[0, 2], 2 * e[1, 2], 2 * e[0, 1]])


@pytest.mark.parametrize("orthotropic", [False, True])
def test_elasticity_energy_identity_and_rigid_modes(orthotropic):
    """Constant strain on a mesh of equal cubes sheared by a symmetric map: every
    element Jacobian is symmetric there, where iron's DXI_DX(mi,ni) index order
    (linear_elasticity_routines.F90:1248) coincides with the chain rule."""
    ne = (3, 2, 2)
    extent = (1.5, 1.0, 1.0)          # equal cubes of side 0.5
    ob, om, od = _mesh(ne, 1, extent, ncomp=3)
    S = np.array([[1.0, 0.2, -0.1], [0.2, 0.9, 0.15], [-0.1, 0.15, 1.1]])
    X = om.coords @ S.T
    E = om.num_elements
    if orthotropic:
        E1, E2, E3 = np.full(E, 12.0), np.full(E, 7.0), np.full(E, 3.0)
        v13, v23, v12 = np.full(E, 0.1), np.full(E, 0.25), np.full(E, 0.3)
    else:
        E1 = E2 = E3 = np.full(E, 10.0)
        v13 = v23 = v12 = np.full(E, 0.3)
    C = fo.linear_elasticity_tensor(E1, E2, E3, v13, v23, v12)
    rp, col = fo.csr_pattern(od)
    xe = X[om.elem_nodes[0] - 1]
    vals = fo.assemble_csr(rp, col, od.element_dofs(), fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs)
    K = fo._csr_scipy(rp, col, vals)
    N = len(X)
    V = np.prod(extent) * np.linalg.det(S)
    rng = np.random.default_rng(5)
    inner = ~_boundary(om.coords, extent)
    for _ in range(4):
        G = rng.standard_normal((3, 3))
        v = np.concatenate([(X @ G.T)[:, c] for c in range(3)])   # separated DOF order
        e = _voigt_strain(G)
        assert v @ (K @ v) == pytest.approx(V * e @ C[0] @ e, rel=1e-11)
        r = K @ v
        for c in range(3):
            assert np.abs(r[c * N:(c + 1) * N][inner]).max() < 1e-11 * np.abs(r).max()
    # rigid motions: translations and infinitesimal rotations (G skew)
    for c in range(3):
        t = np.zeros(3 * N)
        t[c * N:(c + 1) * N] = 1.0
        assert np.abs(K @ t).max() < 1e-11 * abs(K).max()
    W = np.array([[0.0, 1.0, -2.0], [-1.0, 0.0, 0.5], [2.0, -0.5, 0.0]])
    w = np.concatenate([(X @ W.T)[:, c] for c in range(3)])
    assert np.abs(K @ w).max() < 1e-11 * abs(K).max() * np.abs(w).max()


def test_elasticity_orthotropic_blocks_are_distinct():
    """With E1 != E2 != E3 and three Poisson ratios, each shear modulus and each
    normal stiffness appears in a different block: uniaxial strains and single shears
    pick out C11, C22, C33, C44 (g13), C55 (g23) and C66 (g12) one at a time."""
    ne, extent = (2, 2, 2), (1.0, 1.0, 1.0)
    ob, om, od = _mesh(ne, 1, extent, ncomp=3)
    E = om.num_elements
    C = fo.linear_elasticity_tensor(np.full(E, 12.0), np.full(E, 7.0), np.full(E, 3.0), np.full(E, 0.1),
                                    np.full(E, 0.25), np.full(E, 0.3))
    rp, col = fo.csr_pattern(od)
    xe = om.coords[om.elem_nodes[0] - 1]
    K = fo._csr_scipy(rp, col, fo.assemble_csr(rp, col, od.element_dofs(),
                                               fo.linear_elasticity_element_matrices(xe, ob, C), od.num_dofs))
    X = om.coords
    N = len(X)
    C0 = C[0]
    assert len({round(C0[i, i], 9) for i in range(6)}) == 6
    cases = [((0, 0), 0), ((1, 1), 1), ((2, 2), 2), ((0, 2), 3), ((1, 2), 4), ((0, 1), 5)]
    for (i, j), k in cases:
        G = np.zeros((3, 3))
        G[i, j] = 1.0
        G[j, i] = 1.0 if i != j else 1.0
        if i != j:
            G *= 0.5            # engineering shear strain 1
        v = np.concatenate([(X @ G.T)[:, c] for c in range(3)])
        assert v @ (K @ v) == pytest.approx(C0[k, k], rel=1e-12), (i, j)


def test_elasticity_index_order_as_written_on_a_general_affine_mesh(monkeypatch):
    """On a non-symmetric element Jacobian iron's DPHIDX(ns,mi) = sum_ni
    DXI_DX(mi,ni) dphi/dxi_ni (linear_elasticity_routines.F90:1245-1252) is not the
    chain rule: the oracle, which restates it as written, loses the rotation modes,
    and the same oracle with DXI_DX transposed (the chain rule) keeps them.  So the
    oracle's one deviation from the textbook matrix is that index order."""
    ne, extent = (3, 2, 2), (1.5, 1.0, 1.0)
    ob, om, od = _mesh(ne, 1, extent, ncomp=3)
    A = np.array([[1.0, 0.35, -0.2], [0.1, 0.8, 0.25], [-0.15, 0.3, 1.2]])
    X = om.coords @ A.T
    E = om.num_elements
    C = fo.linear_elasticity_tensor(*[np.full(E, v) for v in (10.0, 10.0, 10.0, 0.3, 0.3, 0.3)])
    rp, col = fo.csr_pattern(od)
    xe = X[om.elem_nodes[0] - 1]
    N = len(X)
    W = np.array([[0.0, 1.0, -2.0], [-1.0, 0.0, 0.5], [2.0, -0.5, 0.0]])
    w = np.concatenate([(X @ W.T)[:, c] for c in range(3)])

    def stiffness():
        Ke = fo.linear_elasticity_element_matrices(xe, ob, C)
        return fo._csr_scipy(rp, col, fo.assemble_csr(rp, col, od.element_dofs(), Ke, od.num_dofs))

    K = stiffness()
    assert np.abs(K @ w).max() > 1e-2 * abs(K).max()
    metrics = fo.gauss_metrics
    monkeypatch.setattr(fo, "gauss_metrics", lambda x, b: (lambda m: (m[0], m[1], m[2], np.swapaxes(m[3], -1, -2)))(
        metrics(x, b)))
    Kc = stiffness()
    assert np.abs(Kc @ w).max() < 1e-11 * abs(Kc).max() * np.abs(w).max()
    G = np.array([[0.2, -0.4, 0.1], [0.3, 0.5, -0.2], [0.0, 0.6, -0.1]])
    v = np.concatenate([(X @ G.T)[:, c] for c in range(3)])
    e = _voigt_strain(G)
    assert v @ (Kc @ v) == pytest.approx(np.prod(extent) * np.linalg.det(A) * e @ C[0] @ e, rel=1e-11)
