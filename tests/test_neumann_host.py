"""The Neumann integration matrix (host construction, no GPU) against the oracle's
independent restatement of BoundaryConditions_NeumannIntegrate (fe_oracle.neumann_integrate,
pinned by the unit-traction total and the cantilever tests)."""
import numpy as np
import pytest

from iron_amd import basis, engine, mesh
from oracle import fe_oracle as fo


@pytest.mark.parametrize("order,ne,extent", [(1, (3, 2, 2), (3.0, 1.0, 1.0)), (2, (2, 2, 1), (1.0, 2.0, 0.5))])
def test_integration_matrix_times_point_values_matches_the_oracle(order, ne, extent):
    rng = np.random.default_rng(order)
    b = basis.basis(3, basis.LINEAR_LAGRANGE if order == 1 else basis.QUADRATIC_LAGRANGE)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0, 0, 0])
    X = m.coords
    N1 = len(X)
    face = np.nonzero(np.abs(X[:, 0] - extent[0]) < 1e-12)[0]
    top = np.nonzero(np.abs(X[:, 2] - extent[2]) < 1e-12)[0]
    pv = {int(2 * N1 + i): float(rng.standard_normal()) for i in face}
    pv.update({int(N1 + i): float(rng.standard_normal()) for i in top})
    ring = frozenset(int(2 * N1 + i) for i in top[:2])
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    for io in (frozenset(), ring):
        f_ref = fo.neumann_integrate(om, od, pv, io)
        f = engine.neumann_integrate(m, dm, [b] * 3, pv, integrated_only=io)
        assert np.max(np.abs(f - f_ref)) <= 1e-14 * np.abs(f_ref).max()
    pd = np.array(sorted(pv))
    rp, col, v = engine.neumann_matrix(m, dm, [b] * 3, pd)
    for r in range(dm.num_dofs):
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)
    # unit point values on the end face, the side faces next to it excluded by
    # NEUMANN_INTEGRATED_ONLY DOFs one element in: the rows sum to the face area
    hx = extent[0] / ne[0]
    onb = (np.abs(X[:, 1]) < 1e-12) | (np.abs(X[:, 1] - extent[1]) < 1e-12) | (np.abs(X[:, 2]) < 1e-12) | \
        (np.abs(X[:, 2] - extent[2]) < 1e-12)
    ring = frozenset(int(2 * N1 + i) for i in np.nonzero((np.abs(X[:, 0] - (extent[0] - hx)) < 1e-12) & onb)[0])
    f1 = engine.neumann_integrate(m, dm, [b] * 3, {int(2 * N1 + i): 1.0 for i in face}, integrated_only=ring)
    assert abs(f1.sum() - extent[1] * extent[2]) < 1e-12


@pytest.mark.parametrize("order", [1, 2])
def test_moved_face_nodes(order):
    """Face nodes moved inside the face's plane: the face is the same planar
    rectangle, so unit point values still integrate to its area (the Gauss rule is
    exact for the planar face metric), while the nodal forces change; the host
    matrix built from the moved coordinates matches the oracle's on them."""
    ne, extent = (2, 3, 2), (2.0, 1.0, 1.5)
    b = basis.basis(3, basis.LINEAR_LAGRANGE if order == 1 else basis.QUADRATIC_LAGRANGE)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0, 0, 0])
    X0 = m.coords
    N1 = len(X0)
    face = np.abs(X0[:, 0] - extent[0]) < 1e-12
    inner = face & (X0[:, 1] > 1e-12) & (X0[:, 1] < extent[1] - 1e-12) & (X0[:, 2] > 1e-12) & \
        (X0[:, 2] < extent[2] - 1e-12)
    rng = np.random.default_rng(order)
    X = X0.copy()
    h = np.array(extent) / np.array(ne) / order
    X[inner, 1:] += 0.2 * h[1:] * rng.uniform(-1, 1, (int(inner.sum()), 2))
    pv = {int(2 * N1 + i): 1.0 for i in np.nonzero(face)[0]}
    # the side faces next to the end face are left out by NEUMANN_INTEGRATED_ONLY
    # DOFs one element in
    hx = extent[0] / ne[0]
    onb = (np.abs(X0[:, 1]) < 1e-12) | (np.abs(X0[:, 1] - extent[1]) < 1e-12) | (np.abs(X0[:, 2]) < 1e-12) | \
        (np.abs(X0[:, 2] - extent[2]) < 1e-12)
    ring = frozenset(int(2 * N1 + i) for i in np.nonzero((np.abs(X0[:, 0] - (extent[0] - hx)) < 1e-12) & onb)[0])
    f = engine.neumann_integrate(m, dm, [b] * 3, pv, integrated_only=ring, coords=X)
    f0 = engine.neumann_integrate(m, dm, [b] * 3, pv, integrated_only=ring)
    assert abs(f.sum() - extent[1] * extent[2]) < 1e-12
    assert np.abs(f - f0).max() > 1e-3
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], extent)
    om.coords = X.copy()
    od = fo.DofMap(om, [0, 0, 0])
    q = {d: float(rng.standard_normal()) for d in pv}
    f_ref = fo.neumann_integrate(om, od, q)
    f = engine.neumann_integrate(m, dm, [b] * 3, q, coords=X)
    assert np.abs(f - f_ref).max() <= 1e-14 * np.abs(f_ref).max()
    assert abs(fo.neumann_integrate(om, od, pv, ring).sum() - extent[1] * extent[2]) < 1e-12
