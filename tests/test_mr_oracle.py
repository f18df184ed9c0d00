"""Pins the Mooney-Rivlin oracle (oracle/mr_oracle.py) without a GPU.

* reference state: zero residual;
* the exact Jacobian against central finite differences of the residual;
* homogeneous uniaxial stretch against its closed form;
* the reference's golden cantilever outputs (tests/golden/mr_cantilever_*.json).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import mr_oracle as mo
from oracle.fe_oracle import Basis, regular_mesh

GOLDEN = Path(__file__).parent / "golden"
FORMS = ("deviatoric", "cauchy")


def _problem(n_elem, order, elem_p, form, c1=2.0, c2=6.0, extent=(1.0, 1.0, 1.0)):
    bu = Basis(3, order, order + 1)
    bases = [bu] if elem_p else [bu, Basis(3, 1, order + 1)]
    m = regular_mesh(n_elem, bases, extent)
    E = m.num_elements
    return mo.MRProblem(m, elem_p, np.full(E, c1), np.full(E, c2), formulation=form)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("order,elem_p", [(1, True), (2, False)])
def test_reference_state_residual_vanishes(form, order, elem_p):
    prob = _problem((2, 1, 1), order, elem_p, form, extent=(2.0, 1.0, 1.5))
    # deviatoric: sigma_iso is spherical at F = I and is projected out, so p = 0;
    # cauchy: sigma = 2(c1 + 2 c2) I + p I vanishes for p = -2(c1 + 2 c2).
    p0 = 0.0 if form == "deviatoric" else -2.0 * (2.0 + 2.0 * 6.0)
    R, _ = prob.residual(prob.initial_state(p0))
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("order,elem_p", [(1, True), (2, False)])
def test_jacobian_matches_central_differences(form, order, elem_p):
    rng = np.random.default_rng(3)
    prob = _problem((2, 1, 1), order, elem_p, form, extent=(2.0, 1.0, 1.3))
    prob.c1 = np.array([2.0, 3.0])
    prob.c2 = np.array([6.0, 1.5])
    x = prob.initial_state(-1.0)
    x[:3 * prob.nu] += 0.05 * rng.standard_normal(3 * prob.nu)
    x[3 * prob.nu:] += rng.standard_normal(prob.npd)
    K = prob.jacobian(x).toarray()
    h = 1e-6
    Kfd = np.zeros_like(K)
    for j in range(prob.ndofs):
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        Kfd[:, j] = (prob.residual(xp)[0] - prob.residual(xm)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() < 1e-7 * np.abs(K).max()
    if form == "deviatoric":
        # the deviatoric form is a potential plus a symmetric constraint: K = K^T
        assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()


def _uniaxial_pressure(form, lam, c1, c2):
    B = np.diag([lam * lam, 1.0 / lam, 1.0 / lam])
    i1 = np.trace(B)
    if form == "deviatoric":
        s = 2.0 * (c1 + c2 * i1) * B - 2.0 * c2 * B @ B
        return (s[0, 0] - s[1, 1]) / 3.0
    # cauchy: sigma = 2 c1 B + 2 c2 (I1 B - B^2) + p I with J = 1
    s = 2.0 * c1 * B + 2.0 * c2 * (i1 * B - B @ B)
    return -s[1, 1]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("order,elem_p", [(1, True), (2, False)])
def test_uniaxial_stretch_closed_form(form, order, elem_p):
    """Rollers on x=0, y=0, z=0 and a prescribed x-stretch on x=1: the homogeneous
    isochoric stretch (lam, lam^-1/2, lam^-1/2) with constant p solves the discrete
    problem exactly (linear fields are in the displacement space)."""
    lam, c1, c2 = 1.2, 2.0, 6.0
    prob = _problem((2, 2, 2), order, elem_p, form, c1, c2)
    nu = prob.nu
    X = prob.geo
    fixed = np.zeros(prob.ndofs, bool)
    target = prob.initial_state(0.0)
    for c in range(3):
        fixed[c * nu:(c + 1) * nu] |= X[:, c] == 0.0
    fixed[:nu] |= X[:, 0] == 1.0
    target[:nu][X[:, 0] == 1.0] = lam
    x0 = prob.initial_state(0.0 if form == "deviatoric" else -2.0 * (c1 + 2.0 * c2))
    x, R = mo.solve(prob, x0, fixed, incremented=fixed.copy(), full_value=target, increments=2)
    z, p = prob.split(x)
    exact = X * np.array([lam, lam ** -0.5, lam ** -0.5])
    assert np.abs(z - exact).max() < 1e-10
    assert np.abs(p - _uniaxial_pressure(form, lam, c1, c2)).max() < 1e-8


def _golden_case(name):
    g = json.loads((GOLDEN / name).read_text())
    s = g["setup"]
    bu = Basis(3, 1, 2)
    m = regular_mesh(tuple(s["n_elem"]), [bu], tuple(s["extent"]))
    E = m.num_elements
    prob = mo.MRProblem(m, True, np.full(E, s["c1"]), np.full(E, s["c2"]), s["density"],
                        tuple(s["gravity"]), "geometric", "cauchy")
    fixed = np.zeros(prob.ndofs, bool)
    for c in range(3):
        fixed[c * prob.nu:(c + 1) * prob.nu] = prob.geo[:, 0] == 0.0
    return g, prob, fixed


@pytest.mark.parametrize("name", ["mr_cantilever_python_1x1x1.json", "mr_cantilever_linear_3x2x2.json"])
def test_cantilever_goldens(name):
    """The golden files carry the general-branch (cauchy) solution: deformed nodes and
    DELUDELN to the reference's Newton tolerance; their pressure is half the pressure
    of S = ... + p C^-1 (they were written with S = ... + 2 p C^-1)."""
    g, prob, fixed = _golden_case(name)
    s = g["setup"]
    x, R = mo.solve(prob, prob.initial_state(s["pressure0"]), fixed, increments=s["increments"])
    z, p = prob.split(x)
    nodes = prob.mesh.comp_nodes[0]
    gz = np.array([g["nodes"][str(n)]["dependent"] for n in nodes])
    gr = np.array([g["nodes"][str(n)]["deludeln"] for n in nodes])
    assert np.abs(z - gz).max() < 1e-9
    ru = R[:3 * prob.nu].reshape(3, -1).T
    assert np.abs(ru - gr).max() < 1e-7 * max(1.0, np.abs(gr).max())
    gp = np.array([g["elements"][str(e + 1)]["pressure"] for e in range(prob.mesh.num_elements)])
    assert np.abs(p - 2.0 * gp).max() < 1e-9 * np.abs(gp).max()
    # element connectivity of the golden file equals the generated mesh's
    for e in range(prob.mesh.num_elements):
        assert list(prob.mesh.elem_nodes[0][e]) == g["elements"][str(e + 1)]["nodes"]


def test_deviatoric_differs_from_goldens():
    """The MOONEY_RIVLIN subtype branch does not reproduce the golden files (they predate
    it): the tip of the 3x2x2 cantilever moves by more than 0.1."""
    g, prob, fixed = _golden_case("mr_cantilever_linear_3x2x2.json")
    prob.formulation = "deviatoric"
    s = g["setup"]
    x, _ = mo.solve(prob, prob.initial_state(0.0), fixed, increments=s["increments"])
    z, _ = prob.split(x)
    nodes = prob.mesh.comp_nodes[0]
    gz = np.array([g["nodes"][str(n)]["dependent"] for n in nodes])
    assert np.abs(z - gz).max() > 0.1


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("order,elem_p,geometry", [(1, True, "distorted"), (2, False, "affine")])
def test_homogeneous_state_in_equilibrium_on_non_box_geometry(form, order, elem_p, geometry):
    """On reference geometry that is not a box (interior nodes moved at random on a
    trilinear mesh; a general affine map of a triquadratic one, so that the Gauss rule
    stays exact), any homogeneous isochoric deformation x = F X with constant p has a
    constant stress: the residual vanishes on every displacement DOF off the boundary
    and on every pressure DOF (J = 1), whatever p is."""
    rng = np.random.default_rng(11)
    ne, extent = (3, 2, 2), (1.5, 1.0, 1.2)
    prob = _problem(ne, order, elem_p, form, extent=extent)
    X0 = prob.geo.copy()
    on_b = np.any((np.abs(X0) < 1e-12) | (np.abs(X0 - np.asarray(extent)) < 1e-12), axis=1)
    if geometry == "distorted":
        h = np.asarray(extent) / np.asarray(ne)
        prob.geo = X0 + np.where(on_b[:, None], 0.0, 0.2 * h * rng.uniform(-1, 1, X0.shape))
        assert np.abs(prob.geo - X0).max() > 0.05
    else:
        prob.geo = X0 @ np.array([[1.0, 0.3, -0.2], [0.1, 0.9, 0.25], [-0.15, 0.2, 1.1]]).T
    F = np.array([[1.1, 0.2, 0.0], [-0.1, 0.95, 0.15], [0.05, 0.0, 1.0]])
    F /= np.cbrt(np.linalg.det(F))
    x = prob.initial_state(-1.7)
    z = prob.geo @ F.T
    x[:3 * prob.nu] = np.concatenate([z[:, c] for c in range(3)])
    R, _ = prob.residual(x)
    Ru = R[:3 * prob.nu].reshape(3, -1)
    scale = np.abs(R).max()
    assert scale > 1e-3                                  # boundary rows carry the traction
    assert np.abs(Ru[:, ~on_b]).max() < 1e-11 * scale
    assert np.abs(R[3 * prob.nu:]).max() < 1e-11 * scale


def _simple_shear():
    g = json.loads((GOLDEN / "mr_simple_shear_2x2x2.json").read_text())
    s = g["setup"]
    bu = Basis(3, 2, 3)
    m = regular_mesh(tuple(s["n_elem"]), [bu], tuple(s["extent"]))
    E = m.num_elements
    prob = mo.MRProblem(m, None, np.full(E, s["c1"]), np.full(E, s["c2"]), formulation="nearly_incompressible",
                        kappa=np.full(E, s["kappa"]))
    z = prob.geo[:, 2]
    top, bottom = z == s["extent"][2], z == 0.0
    fixed = np.tile(top | bottom, 3)
    x0 = prob.initial_state()
    value = x0.copy()
    value[:prob.nu][top] += s["shear"] * s["extent"][0]
    return g, prob, fixed, x0, value


def test_nearly_incompressible_reference_state_and_volumetric_stretch():
    """R = 0 in the reference state (S = 2[(c1 + c2 I1) I - c2 C - (c1 I1 + 2 c2 I2)/3 C^-1]
    vanishes at C = I); under a homogeneous volumetric stretch z = s X the Cauchy stress is
    the pure pressure sigma = kappa (J - 1) I (the isochoric part vanishes at C = s^2 I),
    so the element residual is the traction of that pressure."""
    bu = Basis(3, 2, 3)
    m = regular_mesh((1, 1, 1), [bu], (1.0, 1.0, 1.0))
    prob = mo.MRProblem(m, None, np.array([0.7]), np.array([0.3]), formulation="nearly_incompressible",
                        kappa=np.array([50.0]))
    x = prob.initial_state()
    R, _ = prob.residual(x)
    assert np.abs(R).max() < 1e-13
    s = 1.01
    F = np.eye(3) * s
    sig = mo.gauss_nearly_incompressible_cauchy(F[None], np.array([0.7]), np.array([0.3]), np.array([50.0]))[0]
    assert np.allclose(sig, 50.0 * (s ** 3 - 1.0) * np.eye(3), rtol=1e-12, atol=1e-12)


def test_simple_shear_golden():
    """tests/FiniteElasticity/output_shear/SimpleShear.part0: the nearly incompressible
    Mooney-Rivlin box sheared by 0.1 at its top face; deformed nodes and DELUDELN (the
    reactions at the fixed faces, ~1e-9 elsewhere) against the reference's files."""
    g, prob, fixed, x0, value = _simple_shear()
    x, R = mo.solve(prob, x0, fixed, fixed_value=value)
    z, _ = prob.split(x)
    nodes = prob.mesh.comp_nodes[0]
    gz = np.array([g["nodes"][str(n)]["dependent"] for n in nodes])
    gr = np.array([g["nodes"][str(n)]["deludeln"] for n in nodes])
    print("max node difference", np.abs(z - gz).max(), "max DELUDELN difference", np.abs(R.reshape(3, -1).T - gr).max())
    assert np.abs(z - gz).max() < 1e-7
    assert np.abs(R.reshape(3, -1).T - gr).max() < 1e-5 * np.abs(gr).max()
    for e in range(prob.mesh.num_elements):
        assert list(prob.mesh.elem_nodes[0][e]) == g["elements"][str(e + 1)]["nodes"]
