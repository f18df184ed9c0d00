"""FP64 CPU restatement of OpenCMISS-Iron's Mooney-Rivlin finite-elasticity path.

TEST INFRASTRUCTURE ONLY.  This module is part of the oracle: tests/, the smoke
check and bench.py's cpu_baseline leg use it as the checker.  The product package
(iron_amd/) never imports it.

What it restates (paths relative to the reference's root):
  * mr_element_residual  -- FiniteElasticity_FiniteElementResidualEvaluate, Mooney-Rivlin
    branch, finite_elasticity_routines.F90:1669-1760, with
    FINITE_ELASTICITY_GAUSS_STRESS_TENSOR :6985-7058,
    FINITE_ELASTICITY_PUSH_STRESS_TENSOR :1343-1375 and
    FiniteElasticity_GaussDeformationGradientTensor :5148-5212.  Written the way the
    reference writes it: dphi/dz from the deformed metrics, Cauchy stress in Voigt form.
  * gravity_rhs          -- the gravity loading term of the same routine, :3370-3430.
  * mr_element_jacobian  -- the exact derivative of mr_element_residual (the quantity
    iron's finite-difference Jacobian, equations_set_routines.F90:2839-2962,
    approximates).  It is derived independently, in the reference configuration, as
    the Hessian of W = c1 (I1bar - 3) + c2 (I2bar - 3) plus the p J F^-T coupling.
    Iron's own "equations calculated" MR tangent (:649-872, :879-1296) is flagged by
    iron itself as unvalidated (:720) and is not the derivative of its residual; it is
    not restated.
  * solve                -- the load-increment control loop (FIXED_INCREMENTED ramp,
    equations_set_routines.F90:6261-6550; gravity ramp,
    finite_elasticity_routines.F90:13315-13360) around a Newton iteration on
    F(x) = R_int(x) - F_gravity(x) (SNES with the RHS vector b,
    solver_routines.F90:19118), and the post-solve DELUDELN update
    EQUATIONS_SET_NONLINEAR_RHS_UPDATE (equations_set_routines.F90:1919-2052).

Parity pinning (tests/test_mr_oracle.py):
  * R = 0 in the reference state with p = 0 (sigma = 2(c1+3c2)I - 2c2 I is spherical, its
    deviatoric part vanishes, and J = 1).
  * mr_element_jacobian against central finite differences of mr_element_residual.
  * homogeneous uniaxial stretch: lambda_y = lambda_z = lambda^-1/2 and the closed-form
    pressure that makes sigma_yy = 0.
  * the reference's golden .exnode/.exelem files bindings/python/tests/output_cantilever
    (1x1x1 Q1, element pressure, gravity, 2 increments) and
    tests/FiniteElasticity/output_cantilever/linear_3x2x2 (3x2x2 Q1).  They are
    reproduced by the "cauchy" formulation (deformed nodes to 1e-11, DELUDELN to 1e-9)
    with the pressure doubled: the files were written by a version whose second Piola
    tensor carried 2 p C^-1.  The "deviatoric" branch that the current source runs for the
    MOONEY_RIVLIN subtype gives a different solution (tip moves by 0.34) and is pinned by
    the first three checks only.  output_cantilever/cubic_2x1x1 is cubic Hermite and is
    not restated.
  * the NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN subtype (formulation
    "nearly_incompressible", no pressure field): the compressible branch of the residual,
    finite_elasticity_routines.F90:3293-3360, with the second Piola tensor of
    FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR :5308-5320 pushed forward.  Its Jacobian here is
    a central difference of the element residual (iron's own for this subtype is the
    equations' finite-difference Jacobian).  Pinned by R = 0 in the reference state, by
    the homogeneous volumetric stretch closed form and by the golden
    tests/FiniteElasticity/output_shear/SimpleShear.part0 (deformed nodes and DELUDELN).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .fe_oracle import Basis, RegularMesh, gauss_metrics, _invert_full

# constants.F90:305-306 (0-based here)
VOIGT_TO_TENSOR3 = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
TENSOR_TO_VOIGT3 = np.array([[0, 3, 4], [3, 1, 5], [4, 5, 2]])


# ----------------------------------------------------------------------------
# Gauss-point constitutive routines
# ----------------------------------------------------------------------------
def push_stress_tensor(S, F, J):
    """FINITE_ELASTICITY_PUSH_STRESS_TENSOR, finite_elasticity_routines.F90:1343-1375:
    sigma = t(F) S / J with t the Voigt form of S -> F S F^T."""
    t = np.zeros(F.shape[:-2] + (6, 6))
    for j in range(6):
        aj, bj = VOIGT_TO_TENSOR3[j]
        for i in range(6):
            ai, bi = VOIGT_TO_TENSOR3[i]
            t[..., i, j] = F[..., ai, aj] * F[..., bi, bj]
            if j >= 3:
                t[..., i, j] += F[..., ai, bj] * F[..., bi, aj]
    return np.einsum("...ij,...j->...i", t, S) / J[..., None]


def gauss_stress_tensor(F, J, p, c1, c2):
    """FINITE_ELASTICITY_GAUSS_STRESS_TENSOR, Mooney-Rivlin branch,
    finite_elasticity_routines.F90:7017-7058.  Returns the Cauchy stress in Voigt form."""
    Fm = F * (J ** (-1.0 / 3.0))[..., None, None]
    azl = np.einsum("...ki,...kj->...ij", Fm, Fm)
    i1 = azl[..., 0, 0] + azl[..., 1, 1] + azl[..., 2, 2]
    t1 = -2.0 * c2
    t2 = 2.0 * (c1 + i1 * c2)
    S = np.stack([t1 * azl[..., 0, 0] + t2, t1 * azl[..., 1, 1] + t2, t1 * azl[..., 2, 2] + t2,
                  t1 * azl[..., 1, 0], t1 * azl[..., 2, 0], t1 * azl[..., 2, 1]], axis=-1)
    sig = push_stress_tensor(S, Fm, J)
    third = sig[..., :3].sum(-1) / 3.0
    sig[..., :3] += (-third + p)[..., None]
    return sig


def gauss_cauchy_tensor(F, p, c1, c2):
    """FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR, Mooney-Rivlin case,
    finite_elasticity_routines.F90:6036-6068 with the push-forward :6786-6789:
    S = 2 [c1 I + c2 (I1 I - C)] + p C^-1, sigma = F S F^T / det F.  Returns the 3x3
    Cauchy tensor."""
    eye = np.eye(3)
    azl = np.einsum("...ki,...kj->...ij", F, F)
    azu = np.linalg.inv(azl)
    i1 = azl[..., 0, 0] + azl[..., 1, 1] + azl[..., 2, 2]
    S = (2.0 * (c1[..., None, None] * eye + c2[..., None, None] * (i1[..., None, None] * eye - azl))
         + p[..., None, None] * azu)
    J = np.linalg.det(F)
    return np.einsum("...ik,...kl,...jl->...ij", F, S, F) / J[..., None, None]


def gauss_nearly_incompressible_cauchy(F, c1, c2, kappa):
    """FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR, NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN case,
    finite_elasticity_routines.F90:5308-5320, pushed forward as for the other laws:
    W^v(J) = kappa (J - 1)^2 / 2, WV' = kappa (J - 1),
    S = 2 J^(-2/3) [(c1 + c2 I1) I - c2 C - (c1 I1 + 2 c2 I2 - 1.5 WV' J^(5/3)) / 3 C^-1],
    sigma = F S F^T / J, with C = F^T F, I1 = tr C, I2 = (I1^2 - tr C^2) / 2, J = det F.
    Returns the 3x3 Cauchy tensor."""
    eye = np.eye(3)
    azl = np.einsum("...ki,...kj->...ij", F, F)
    azu = np.linalg.inv(azl)
    J = np.linalg.det(F)
    i1 = azl[..., 0, 0] + azl[..., 1, 1] + azl[..., 2, 2]
    sq = np.einsum("...ij,...jk->...ik", azl, azl)
    i2 = 0.5 * (i1 ** 2 - sq[..., 0, 0] - sq[..., 1, 1] - sq[..., 2, 2])
    wv = kappa * (J - 1.0)
    coef = (c1 * i1 + 2.0 * c2 * i2 - 1.5 * wv * J ** (5.0 / 3.0)) / 3.0
    S = 2.0 * (J ** (-2.0 / 3.0))[..., None, None] * ((c1 + c2 * i1)[..., None, None] * eye
                                                     - c2[..., None, None] * azl - coef[..., None, None] * azu)
    return np.einsum("...ik,...kl,...jl->...ij", F, S, F) / J[..., None, None]


# ----------------------------------------------------------------------------
# Element residual, gravity vector, exact Jacobian
# ----------------------------------------------------------------------------
def _pressure_at_gauss(pe, bp: Basis | None):
    if bp is None:
        return pe[:, None]
    return np.einsum("gq,eq->eg", bp.value_table(), pe)


def mr_element_residual(geo_e, dep_e, pe, bu: Basis, bp: Basis | None, c1, c2,
                        formulation="deviatoric", kappa=None):
    """FiniteElasticity_FiniteElementResidualEvaluate, finite_elasticity_routines.F90.

    formulation "deviatoric": the MOONEY_RIVLIN subtype branch :1669-1756 (isochoric
    second Piola tensor pushed forward, deviatoric part plus p).
    formulation "cauchy": the general branch taken by EQUATIONS_SET_NO_SUBTYPE,
    :1824-2000, with the Mooney-Rivlin law of FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR
    (:6036-6068): S = 2[c1 I + c2 (I1 I - C)] + p C^-1.
    formulation "nearly_incompressible": the compressible branch :3293-3360 with
    gauss_nearly_incompressible_cauchy (kappa [E]); no pressure DOFs (pe, bp unused),
    the result is [E, 3*nn].

    geo_e, dep_e: [E, nn, 3] geometric and deformed nodal coordinates; pe: [E, npn]
    nodal pressures (bp = pressure basis) or [E] element pressures (bp = None);
    c1, c2: [E].  Returns [E, 3*nn + npn] in element-DOF order (component-major:
    x over the element nodes, then y, z, then pressure)."""
    E, nn, _ = geo_e.shape
    dN = bu.dxi_table()                      # [g, nn, 3]
    w = bu.gauss_weights
    _, _, jx, dxi_dx = gauss_metrics(geo_e, bu)
    dz, _, jz, dxi_dz = gauss_metrics(dep_e, bu)
    # DPHIDZ(mh,ns) = sum_mi DXI_DX(mi,mh) dphi_ns/dxi_mi  (:1690-1705)
    dphidz = np.einsum("egim,gni->egmn", dxi_dz, dN)
    # dZ/dNu = dZ/dXi . dXi/dNu with the material frame = the global frame (:5171-5175)
    F = np.einsum("egij,egjk->egik", dz, dxi_dx)
    jznu = jz / jx
    cc1 = np.broadcast_to(c1[:, None], jz.shape)
    cc2 = np.broadcast_to(c2[:, None], jz.shape)
    if formulation == "nearly_incompressible":
        # sum_g w Jxxi Jznu sigma_ck dphi_a/dz_k (:3345-3351)
        T = gauss_nearly_incompressible_cauchy(F, cc1, cc2, np.broadcast_to(kappa[:, None], jz.shape))
        T = T * (jz * w)[..., None, None]
        return np.einsum("egkn,egkm->emn", dphidz, T).reshape(E, 3 * nn)
    p = _pressure_at_gauss(pe, bp) * np.ones_like(jz)
    if formulation == "deviatoric":
        sig = gauss_stress_tensor(F, jznu, p, cc1, cc2)
        T = sig[..., TENSOR_TO_VOIGT3]
    elif formulation == "cauchy":
        T = gauss_cauchy_tensor(F, p, cc1, cc2)
    else:
        raise ValueError(formulation)
    T = T * (jz * w)[..., None, None]                            # JGW_CAUCHY_TENSOR
    ru = np.einsum("egkn,egkm->emn", dphidz, T).reshape(E, 3 * nn)
    tt = jx * w * (jznu - 1.0)                                   # :1738-1752
    if bp is None:
        rp = tt.sum(axis=1)[:, None]
    else:
        rp = np.einsum("eg,gq->eq", tt, bp.value_table())
    return np.concatenate([ru, rp], axis=1)


def gravity_rhs(geo_e, dep_e, bu: Basis, density, gravity, jacobian="geometric"):
    """Gravity loading term, finite_elasticity_routines.F90:3370-3430:
    f_(c,m) += rho g_c phi_m w J, component-major [E, 3*nn].  ``jacobian`` selects the
    volume factor J (the tests decide against the golden files which one the
    reference's outputs carry): "geometric" = reference-configuration Jacobian,
    "deformed_last" = the deformed Jacobian of the last Gauss point of the element."""
    E, nn, _ = geo_e.shape
    N = bu.value_table()
    w = bu.gauss_weights
    _, _, jx, _ = gauss_metrics(geo_e, bu)
    if jacobian == "geometric":
        J = jx
    else:
        _, _, jz, _ = gauss_metrics(dep_e, bu)
        J = np.repeat(jz[:, -1:], jz.shape[1], axis=1)
    rho = np.broadcast_to(np.asarray(density, dtype=np.float64), (E,))
    m = np.einsum("gn,eg->en", N, J * w) * rho[:, None]
    return np.concatenate([m * gravity[c] for c in range(3)], axis=1)


def _w_iso_first_second(F, c1, c2):
    """First and second derivatives of W = c1 (I1bar-3) + c2 (I2bar-3) with respect to F
    (I1bar = J^-2/3 I1, I2bar = J^-4/3 I2).  Returns P [...,3,3] and A [...,3,3,3,3]
    with A[i,K,j,L] = d2W / dF_iK dF_jL."""
    eye = np.eye(3)
    J = np.linalg.det(F)
    G = np.swapaxes(np.linalg.inv(F), -1, -2)        # F^-T
    I1 = np.einsum("...ij,...ij->...", F, F)
    C = np.einsum("...ki,...kj->...ij", F, F)
    B = np.einsum("...ik,...jk->...ij", F, F)
    I2 = 0.5 * (I1 * I1 - np.einsum("...ij,...ij->...", C, C))
    a = J ** (-2.0 / 3.0)
    b = J ** (-4.0 / 3.0)
    FC = np.einsum("...ik,...kj->...ij", F, C)
    I1F = 2.0 * F
    I2F = 2.0 * (I1[..., None, None] * F - FC)
    Ib1F = a[..., None, None] * (I1F - (2.0 / 3.0) * I1[..., None, None] * G)
    Ib2F = b[..., None, None] * (I2F - (4.0 / 3.0) * I2[..., None, None] * G)
    P = c1[..., None, None] * Ib1F + c2[..., None, None] * Ib2F

    dd = np.einsum("ij,KL->iKjL", eye, eye)
    GtG = np.einsum("...iL,...jK->...iKjL", G, G)     # -dG_iK/dF_jL
    ex = lambda X, Y: np.einsum("...iK,...jL->...iKjL", X, Y)
    e4 = lambda x: x[..., None, None, None, None]
    Ib1FF = (-(2.0 / 3.0) * e4(a) * ex(I1F - (2.0 / 3.0) * I1[..., None, None] * G, G)
             + e4(a) * (2.0 * dd - (2.0 / 3.0) * ex(G, I1F) + (2.0 / 3.0) * e4(I1) * GtG))
    I2FF = 2.0 * (2.0 * ex(F, F) + e4(I1) * dd
                  - np.einsum("ij,...LK->...iKjL", eye, C)
                  - np.einsum("...iL,...jK->...iKjL", F, F)
                  - np.einsum("...ij,KL->...iKjL", B, eye))
    Ib2FF = (-(4.0 / 3.0) * e4(b) * ex(I2F - (4.0 / 3.0) * I2[..., None, None] * G, G)
             + e4(b) * (I2FF - (4.0 / 3.0) * ex(G, I2F) + (4.0 / 3.0) * e4(I2) * GtG))
    A = e4(c1) * Ib1FF + e4(c2) * Ib2FF
    return P, A, J, G


def _cauchy_first_second(F, p, c1, c2):
    """P = F S and A[i,K,j,L] = dP_iK/dF_jL for S = 2[c1 I + c2 (I1 I - C)] + p C^-1:
    A = delta_ij S_LK + 2 c2 (2 F_iK F_jL - F_iL F_jK - B_ij delta_KL)
        - p (G_iL G_jK + delta_ij Cinv_LK),  G = F^-T;  dP/dp = G."""
    eye = np.eye(3)
    C = np.einsum("...ki,...kj->...ij", F, F)
    B = np.einsum("...ik,...jk->...ij", F, F)
    Ci = np.linalg.inv(C)
    G = np.swapaxes(np.linalg.inv(F), -1, -2)
    I1 = C[..., 0, 0] + C[..., 1, 1] + C[..., 2, 2]
    S = (2.0 * (c1[..., None, None] * eye + c2[..., None, None] * (I1[..., None, None] * eye - C))
         + p[..., None, None] * Ci)
    P = np.einsum("...iM,...MK->...iK", F, S)
    e4 = lambda x: x[..., None, None, None, None]
    A = (np.einsum("ij,...LK->...iKjL", eye, S)
         + 2.0 * e4(c2) * (2.0 * np.einsum("...iK,...jL->...iKjL", F, F)
                           - np.einsum("...iL,...jK->...iKjL", F, F)
                           - np.einsum("...ij,KL->...iKjL", B, eye))
         - e4(p) * (np.einsum("...iL,...jK->...iKjL", G, G) + np.einsum("ij,...LK->...iKjL", eye, Ci)))
    return P, A, G


def mr_element_jacobian(geo_e, dep_e, pe, bu: Basis, bp: Basis | None, c1, c2,
                        formulation="deviatoric"):
    """d(mr_element_residual)/d(element DOFs) [E, n, n], n = 3*nn + npn.

    In the reference configuration the residual reads
    R_(i,a) = sum_g w Jx dphi_a/dX_K P_iK,   R_q = sum_g w Jx (J - 1) psi_q,
    with P = dW/dF + p J F^-T (deviatoric) or P = F S (cauchy), so
    K_uu = w Jx GX A GX (A = dP/dF), K_up = w Jx GX dP/dp psi, K_pu = w Jx psi J F^-T GX,
    K_pp = 0.  For the deviatoric form dP/dp = J F^-T and K_pu = K_up^T."""
    E, nn, _ = geo_e.shape
    dN = bu.dxi_table()
    w = bu.gauss_weights
    _, _, jx, dxi_dx = gauss_metrics(geo_e, bu)
    dz, _, jz, _ = gauss_metrics(dep_e, bu)
    F = np.einsum("egij,egjk->egik", dz, dxi_dx)
    GX = np.einsum("gnm,egmK->egnK", dN, dxi_dx)          # dphi/dX
    p = _pressure_at_gauss(pe, bp) * np.ones_like(jz)
    cc1 = np.broadcast_to(c1[:, None], jz.shape)
    cc2 = np.broadcast_to(c2[:, None], jz.shape)
    if formulation == "deviatoric":
        _, A, J, G = _w_iso_first_second(F, cc1, cc2)
        JG = J[..., None, None] * G
        A = A + p[..., None, None, None, None] * (np.einsum("...jL,...iK->...iKjL", JG, G)
                                                  - J[..., None, None, None, None]
                                                  * np.einsum("...iL,...jK->...iKjL", G, G))
        dPdp = JG
    elif formulation == "cauchy":
        _, A, G = _cauchy_first_second(F, p, cc1, cc2)
        J = np.linalg.det(F)
        JG = J[..., None, None] * G
        dPdp = G
    else:
        raise ValueError(formulation)
    wj = jx * w
    Kuu = np.einsum("egaK,egiKjL,egbL,eg->eiajb", GX, A, GX, wj).reshape(E, 3 * nn, 3 * nn)
    if bp is None:
        psi = np.ones((len(w), 1))
    else:
        psi = bp.value_table()
    Kup = np.einsum("egaK,egiK,gq,eg->eiaq", GX, dPdp, psi, wj).reshape(E, 3 * nn, psi.shape[1])
    Kpu = np.einsum("egaK,egiK,gq,eg->eiaq", GX, JG, psi, wj).reshape(E, 3 * nn, psi.shape[1])
    npn = psi.shape[1]
    K = np.zeros((E, 3 * nn + npn, 3 * nn + npn))
    K[:, :3 * nn, :3 * nn] = Kuu
    K[:, :3 * nn, 3 * nn:] = Kup
    K[:, 3 * nn:, :3 * nn] = np.swapaxes(Kpu, 1, 2)
    return K


# ----------------------------------------------------------------------------
# Global problem: DOF layout, assembly, load-increment Newton
# ----------------------------------------------------------------------------
@dataclass
class MRProblem:
    """A regular mesh with displacement basis ``mesh.bases[0]`` and either a node-based
    pressure on mesh component 1 (``p_element_based`` False) or one pressure per element.

    DOF layout (SEPARATED order, field_routines.F90:10056): x over the displacement
    mesh nodes, then y, then z, then the pressure DOFs (pressure-mesh nodes or
    elements in element order)."""
    mesh: RegularMesh
    p_element_based: bool
    c1: np.ndarray
    c2: np.ndarray
    density: float = 0.0
    gravity: tuple | None = None
    gravity_jacobian: str = "geometric"
    formulation: str = "deviatoric"
    kappa: np.ndarray | None = None   # nearly_incompressible: bulk penalty per element

    def __post_init__(self):
        m = self.mesh
        self.bu = m.bases[0]
        self.bp = None if (self.p_element_based or self.p_element_based is None) else m.bases[1]
        self.nu = len(m.comp_nodes[0])
        lut = np.full(int(m.comp_nodes[0].max()) + 1, -1, dtype=np.int64)
        lut[m.comp_nodes[0]] = np.arange(self.nu)
        self.en_u = lut[m.elem_nodes[0]]
        if self.p_element_based is None:   # no pressure field (nearly incompressible)
            self.npd = 0
            self.en_p = np.zeros((m.num_elements, 0), dtype=np.int64)
        elif self.p_element_based:
            self.npd = m.num_elements
            self.en_p = np.arange(m.num_elements)[:, None]
        else:
            pn = m.comp_nodes[1]
            lutp = np.full(int(pn.max()) + 1, -1, dtype=np.int64)
            lutp[pn] = np.arange(len(pn))
            self.npd = len(pn)
            self.en_p = lutp[m.elem_nodes[1]]
        self.ndofs = 3 * self.nu + self.npd
        self.edofs = np.concatenate([self.en_u + c * self.nu for c in range(3)]
                                    + [self.en_p + 3 * self.nu], axis=1)
        self.geo = np.zeros((self.nu, 3))
        self.geo[:, :m.dim] = m.coords

    def split(self, x):
        z = x[:3 * self.nu].reshape(3, self.nu).T
        return z, x[3 * self.nu:]

    def element_arrays(self, x):
        z, p = self.split(x)
        geo_e = self.geo[self.en_u]
        dep_e = z[self.en_u]
        if self.npd == 0:
            pe = None
        else:
            pe = p[self.en_p[:, 0]] if self.p_element_based else p[self.en_p]
        return geo_e, dep_e, pe

    def residual(self, x, load=1.0):
        """(R_int, F_ext) as global vectors: R_int the assembled nonlinear residual, F_ext
        the assembled gravity vector scaled by the current load fraction."""
        geo_e, dep_e, pe = self.element_arrays(x)
        re = mr_element_residual(geo_e, dep_e, pe, self.bu, self.bp, self.c1, self.c2, self.formulation,
                                 self.kappa)
        R = np.zeros(self.ndofs)
        np.add.at(R, self.edofs, re)
        f = np.zeros(self.ndofs)
        if self.gravity is not None and self.density != 0.0:
            g = np.asarray(self.gravity, dtype=np.float64) * load
            fe = gravity_rhs(geo_e, dep_e, self.bu, self.density, g, self.gravity_jacobian)
            np.add.at(f, self.edofs[:, :fe.shape[1]], fe)
        return R, f

    def element_jacobian_fd(self, x, h=1e-7):
        """Central difference of the element residual in each element DOF (all elements
        at once): [E, n, n]."""
        geo_e, dep_e, pe = self.element_arrays(x)
        E, nn, _ = dep_e.shape
        n = 3 * nn
        K = np.zeros((E, n, n))
        for k in range(n):
            c, a = divmod(k, nn)
            dp, dm = dep_e.copy(), dep_e.copy()
            dp[:, a, c] += h
            dm[:, a, c] -= h
            rp = mr_element_residual(geo_e, dp, pe, self.bu, self.bp, self.c1, self.c2, self.formulation, self.kappa)
            rm = mr_element_residual(geo_e, dm, pe, self.bu, self.bp, self.c1, self.c2, self.formulation, self.kappa)
            K[:, :, k] = (rp - rm) / (2.0 * h)
        return K

    def jacobian(self, x):
        if self.formulation == "nearly_incompressible":
            ke = self.element_jacobian_fd(x)
        else:
            geo_e, dep_e, pe = self.element_arrays(x)
            ke = mr_element_jacobian(geo_e, dep_e, pe, self.bu, self.bp, self.c1, self.c2, self.formulation)
        n = self.edofs.shape[1]
        rows = np.repeat(self.edofs, n, axis=1).ravel()
        cols = np.tile(self.edofs, (1, n)).ravel()
        return sp.csr_matrix((ke.ravel(), (rows, cols)), shape=(self.ndofs, self.ndofs))

    def initial_state(self, pressure0=0.0):
        x = np.concatenate([self.geo[:, c] for c in range(3)] + [np.full(self.npd, pressure0)])
        return x


def newton(prob: MRProblem, x, free, load=1.0, rtol=1e-13, atol=1e-11, maxit=50):
    """Full Newton on the free DOFs with a backtracking safeguard; converges far below
    the reference's SNES tolerances so the root, not the iteration path, is compared."""
    fr = np.nonzero(free)[0]
    R, f = prob.residual(x, load)
    F = (R - f)[fr]
    f0 = np.linalg.norm(F)
    it = 0
    while np.linalg.norm(F) > max(atol, rtol * max(f0, 1.0)) and it < maxit:
        K = prob.jacobian(x)[fr][:, fr].tocsc()
        dx = spla.spsolve(K, -F)
        step = 1.0
        fn = np.linalg.norm(F)
        while True:
            xt = x.copy()
            xt[fr] += step * dx
            Rt, ft = prob.residual(xt, load)
            Ft = (Rt - ft)[fr]
            if np.all(np.isfinite(Ft)) and (np.linalg.norm(Ft) <= (1.0 - 1e-4 * step) * fn or step < 1e-4):
                break
            step *= 0.5
        x, F = xt, Ft
        it += 1
    if np.linalg.norm(F) > max(atol, rtol * max(f0, 1.0)) * 10:
        raise RuntimeError(f"oracle Newton did not converge: |F|={np.linalg.norm(F):.3e}")
    return x, it


def solve(prob: MRProblem, x0, fixed, fixed_value=None, incremented=None, full_value=None,
          increments=1, **kw):
    """Load-increment loop: iteration i of N sets each FIXED_INCREMENTED DOF to
    current + (full - current)/(N - i + 1) (equations_set_routines.F90:6340-6350) and
    scales gravity by i/N when N > 1 (finite_elasticity_routines.F90:13333-13350), then
    runs Newton.  Returns the final state and R_int there (the DELUDELN values of
    EQUATIONS_SET_NONLINEAR_RHS_UPDATE for RHS-free DOFs)."""
    x = x0.copy()
    if fixed_value is not None:
        x[fixed] = fixed_value[fixed]
    free = ~fixed
    for i in range(1, increments + 1):
        if incremented is not None and incremented.any():
            cur = x[incremented]
            x[incremented] = cur + (full_value[incremented] - cur) / (increments - i + 1)
        load = i / increments if increments > 1 else 1.0
        x, _ = newton(prob, x, free, load, **kw)
    R, _ = prob.residual(x, 1.0)
    return x, R
