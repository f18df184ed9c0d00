"""NeumannIntegrate's f = N q on the device (engine.NeumannIntegrator, CSR product
through mife_spmv) against the oracle's restatement, Q1 and Q2, with and without
NEUMANN_INTEGRATED_ONLY DOFs.  FP64 to 1e-14 of the largest entry."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import fe_oracle as fo  # noqa: E402


@pytest.mark.parametrize("order,ne,extent", [(1, (4, 2, 3), (3.0, 1.0, 1.5)), (2, (2, 3, 2), (1.0, 2.0, 0.5))])
def test_device_integration_matches_the_oracle(order, ne, extent):
    from iron_amd import basis, engine, mesh
    print(f"neumann order={order} ne={ne}", flush=True)
    rng = np.random.default_rng(10 + order)
    b = basis.basis(3, basis.LINEAR_LAGRANGE if order == 1 else basis.QUADRATIC_LAGRANGE)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0, 0, 0])
    X = m.coords
    N1 = len(X)
    face = np.nonzero(np.abs(X[:, 0] - extent[0]) < 1e-12)[0]
    top = np.nonzero(np.abs(X[:, 2] - extent[2]) < 1e-12)[0]
    pv = {int(2 * N1 + i): float(rng.standard_normal()) for i in face}
    pv.update({int(i): float(rng.standard_normal()) for i in top})
    ob = fo.Basis(3, order, order + 1)
    om = fo.regular_mesh(ne, [ob], extent)
    od = fo.DofMap(om, [0, 0, 0])
    dev = torch.device("cuda:0")
    pd = np.array(sorted(pv))
    q = torch.as_tensor(np.array([pv[int(d)] for d in pd]), device=dev)
    for io in (frozenset(), frozenset(int(i) for i in top[::3])):
        ni = engine.NeumannIntegrator(m, dm, [b] * 3, pd, dev, integrated_only=io)
        f = ni.integrate(q).cpu().numpy()
        f_ref = fo.neumann_integrate(om, od, pv, io)
        assert np.max(np.abs(f - f_ref)) <= 1e-14 * np.abs(f_ref).max()
