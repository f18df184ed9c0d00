"""Mooney-Rivlin finite elasticity on the GPU: element residual/Jacobian kernel,
load-increment control loop and Newton iteration.

Reference call stack (paths relative to the reference's root):
  Problem_Solve -> load-increment control loop (CONTROL_LOOP_LOAD_INCREMENT):
    EQUATIONS_SET_BOUNDARY_CONDITIONS_INCREMENT (equations_set_routines.F90:6261-6550)
    FINITE_ELASTICITY_LOAD_INCREMENT_APPLY (finite_elasticity_routines.F90:13315-13360)
    SOLVER_NEWTON_LINESEARCH_SOLVE -> SNESSolve(b = gravity RHS) (solver_routines.F90:19062-19220)
      residual: FiniteElasticity_FiniteElementResidualEvaluate (:1382-3506) + ElementAdd
      Jacobian: FiniteElasticity_FiniteElementJacobianEvaluate (:879-1296) or FD
      linear solve: KSP on the free rows
    EQUATIONS_SET_NONLINEAR_RHS_UPDATE (equations_set_routines.F90:1919-2052)

Here the residual, the exact Jacobian and the gravity vector come from one HIP
kernel (mife_mr_assemble), the Newton system is the free-row reduction of the
Jacobian in SELL-64 storage, and the Newton step is solved with preconditioned
MINRES (symmetric deviatoric form) or GMRES (general form) whose diagonal
preconditioner uses 1/|a_ii| on displacement rows and the element-lumped
pressure mass scaled by 1/(2(c1+c2)) on pressure rows.  The line search is
backtracking with the Armijo factor 1e-4 (SNES's default alpha, solver_routines.F90:18942-18991).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, fem
from ._lib import CMFEError, MrDesc, MrPlan, call, ptr
from .basis import LagrangeBasis
from .mesh import RegularMesh, structured_colours

DEVIATORIC = _lib.MIFE_MR_DEVIATORIC
CAUCHY = _lib.MIFE_MR_CAUCHY
NEARLY_INCOMPRESSIBLE = _lib.MIFE_MR_NEARLY_INCOMPRESSIBLE


class MRSystem:
    """Device structures of a Mooney-Rivlin problem on a regular mesh.

    Displacement (and geometry) on mesh component 0 (``rm.bases[0]``); pressure
    either node based on mesh component ``p_mesh_comp`` or element based
    (``p_mesh_comp=None``), and absent for the NEARLY_INCOMPRESSIBLE formulation
    (``p_mesh_comp`` ignored).  DOF layout (separated order): x over the
    displacement nodes, then y, then z, then the pressure DOFs.  ``coords`` are the
    reference node coordinates on mesh component 0 (the geometric field's values;
    default the generated mesh's)."""

    def __init__(self, rm: RegularMesh, p_mesh_comp: int | None, formulation: int = DEVIATORIC, device=None,
                 coords: np.ndarray | None = None):
        if rm.dim != 3:
            raise CMFEError(1, "finite elasticity is implemented for 3-D meshes")
        dev = fem._dev(device)
        self.device = dev
        self.rm = rm
        self.formulation = int(formulation)
        bu: LagrangeBasis = rm.bases[0]
        self.bu = bu
        self.nu = rm.num_nodes(0)
        en_u = rm.elem_nodes[0]
        E = rm.num_elements
        if self.formulation == NEARLY_INCOMPRESSIBLE:
            self.bp = None
            self.npd = 0
            en_p = np.zeros((E, 0), dtype=np.int32)
            self.n_pressure = 0
        elif p_mesh_comp is None:
            self.bp = None
            self.npd = E
            en_p = np.arange(E, dtype=np.int32)[:, None]
            self.n_pressure = 1
        else:
            bp: LagrangeBasis = rm.bases[p_mesh_comp]
            if bp.gauss_xi != bu.gauss_xi:
                raise CMFEError(1, "the pressure basis must use the displacement basis' Gauss scheme")
            self.bp = bp
            self.npd = rm.num_nodes(p_mesh_comp)
            en_p = rm.elem_nodes[p_mesh_comp]
            self.n_pressure = bp.num_nodes
        self.ndofs = 3 * self.nu + self.npd
        self.elem_dofs = np.ascontiguousarray(
            np.concatenate([en_u + c * self.nu for c in range(3)] + [en_p + 3 * self.nu], axis=1).astype(np.int32))
        self.en_p = en_p
        self.csr = fem.CSRStructure(self.elem_dofs, self.ndofs, dev, colours=structured_colours(rm))
        self.db = fem.DeviceBasis(bu, dev)
        self.Np = None
        if self.bp is not None:
            self.Np = torch.as_tensor(np.ascontiguousarray(self.bp.N), dtype=torch.float64, device=dev)
        self.X = np.ascontiguousarray(rm.coords if coords is None else coords, dtype=np.float64)
        if self.X.shape != (self.nu, 3):
            raise CMFEError(1, "reference coordinates need one row of 3 values per displacement node")
        self.coords = torch.as_tensor(self.X, device=dev)
        self.elem_nodes = torch.as_tensor(np.ascontiguousarray(en_u, dtype=np.int32), device=dev)
        self.desc = MrDesc(C.pointer(self.db.desc), self.n_pressure, ptr(self.Np), self.formulation)
        # staged scatter: one element launch into a stage buffer, then fixed-order
        # sums by a plan built once from the scatter map
        cs = self.csr
        ne = cs.n_edof
        E = rm.num_elements
        self._plan_arrays = [torch.empty(cs.nnz + 1, dtype=torch.int32, device=dev),
                             torch.empty(max(E * ne * ne, 1), dtype=torch.int32, device=dev),
                             torch.empty(self.ndofs + 1, dtype=torch.int32, device=dev),
                             torch.empty(max(E * ne, 1), dtype=torch.int32, device=dev)]
        wb = int(_lib.lib().mife_mr_plan_work_bytes(E, ne, cs.nnz, self.ndofs))
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        call("mife_mr_plan_build", E, ne, ptr(cs.pos), cs.nnz, ptr(cs.elem_dofs), self.ndofs,
             *[ptr(a) for a in self._plan_arrays], ptr(work), wb, fem._stream())
        del work
        self.plan = MrPlan(cs.nnz, self.ndofs, 0, *[ptr(a) for a in self._plan_arrays])
        sb = int(_lib.lib().mife_mr_stage_bytes(C.byref(self.desc), E))
        self.stage = torch.zeros(max(sb // 8, 1), dtype=torch.float64, device=dev)

    def assemble(self, state, material, body_force=None, residual=None, jac=None, rhs=None, atomic=False,
                 staged=True):
        """Accumulate the residual, the Jacobian (equations-CSR values) and the gravity
        vector at ``state`` into the given (zeroed) device arrays.  Staged (default):
        one element launch and sums in element order; else one launch per colour
        (``atomic``: one launch with atomic adds)."""
        cs = self.csr
        bf = None
        if body_force is not None:
            bf = (C.c_double * 3)(*[float(v) for v in body_force])
        if staged and not atomic:
            call("mife_mr_assemble_staged", C.byref(self.desc), self.rm.num_elements, ptr(self.coords),
                 ptr(self.elem_nodes), ptr(cs.elem_dofs), ptr(state), ptr(material),
                 C.cast(bf, C.c_void_p) if bf is not None else None, C.byref(self.plan), ptr(self.stage),
                 ptr(residual), ptr(jac), ptr(rhs), fem._stream())
            return
        call("mife_mr_assemble", C.byref(self.desc), self.rm.num_elements, ptr(self.coords), ptr(self.elem_nodes),
             ptr(cs.elem_dofs), ptr(state), ptr(material), C.cast(bf, C.c_void_p) if bf is not None else None,
             ptr(cs.pos) if jac is not None else None, None if atomic else ptr(cs.order),
             None if atomic else cs.colour_ptr.ctypes.data, cs.n_colours, ptr(residual), ptr(jac), ptr(rhs),
             fem._stream())

    def assemble_colour(self, c: int, state, material, body_force=None, residual=None, rhs=None):
        """As assemble, residual and gravity vector only, over the elements of colour c
        (no two of which share a DOF)."""
        cs = self.csr
        b0, b1 = int(cs.colour_ptr[c]), int(cs.colour_ptr[c + 1])
        cp = np.array([0, b1 - b0], dtype=np.int32)
        bf = None
        if body_force is not None:
            bf = (C.c_double * 3)(*[float(v) for v in body_force])
        call("mife_mr_assemble", C.byref(self.desc), self.rm.num_elements, ptr(self.coords), ptr(self.elem_nodes),
             ptr(cs.elem_dofs), ptr(state), ptr(material), C.cast(bf, C.c_void_p) if bf is not None else None,
             None, ptr(cs.order) + 4 * b0, cp.ctypes.data, 1, ptr(residual), None, ptr(rhs), fem._stream())

    def pressure_mass(self, mu_e: np.ndarray) -> np.ndarray:
        """Lumped sum_e (1/mu_e) int psi_q dV over the pressure DOFs ([ndofs], zero on
        displacement DOFs): the diagonal Schur-complement estimate."""
        if self.npd == 0:
            return np.zeros(self.ndofs)
        bu = self.bu
        X = self.X[self.rm.elem_nodes[0]]                          # [E, nn, 3]
        J = np.einsum("ead,gaj->egdj", X, bu.dN)
        wj = np.abs(np.linalg.det(J)) * bu.weights[None, :]         # [E, ng]
        psi = np.ones((bu.num_gauss, 1)) if self.bp is None else self.bp.N
        m = np.einsum("eg,gq->eq", wj, psi) / mu_e[:, None]
        out = np.zeros(self.ndofs)
        np.add.at(out, 3 * self.nu + self.en_p, m)
        return out


@dataclass
class NewtonOptions:
    """SNES settings (iron defaults, solver_routines.F90:17960-18030)."""
    atol: float = 1e-10
    rtol: float = 1e-5
    stol: float = 1e-5
    maxit: int = 50
    linear_rtol: float = 1e-5       # or "ew": Eisenstat-Walker forcing terms
    linear_maxit: int = 100000
    linear: str = "auto"          # "minres", "gmres" or "auto" (by formulation)
    gmres_restart: int = 100
    # values of the Newton matrix inside MINRES: "f32" (FP32 values, FP64 vectors: the
    # Newton residual stays FP64, so only the linear solve's operator is rounded) or "f64"
    minres_values: str = "f32"
    # Jacobian: "exact" (analytic consistent tangent) or "fd" (iron's one-sided element
    # finite differences, step (1 + ||x||_2) fd_step)
    jacobian: str = "exact"
    fd_step: float = 1e-6
    # line search on f = ||F||^2 / 2 along the Newton step (iron's
    # SOLVER_NEWTON_LINESEARCH_*, solver_routines.F90:18826-18845): "basic" (full step),
    # "linear" (one secant step to the critical point of f), "quadratic" or "cubic"
    # (backtracking with quadratic / cubic models and sufficient decrease alpha)
    linesearch: str = "cubic"
    # convergence test: "petsc" (atol / rtol on ||F||, stol on the step) or "energy"
    # (iron's SOLVER_NEWTON_CONVERGENCE_ENERGY_NORM, problem_routines.F90:4232-4300:
    # |step . F(new)| relative to its value after the first step, below atol)
    convergence: str = "petsc"
    ls_alpha: float = 1e-4
    ls_maxit: int = 40
    ls_minlambda: float = 1e-12


def backtrack_lambda(order: str, lam: float, f0: float, slope: float, f_lam: float,
                     lam_prev: float | None = None, f_prev: float | None = None) -> float:
    """Next step length of a backtracking line search on f(lambda) = ||F(x + lambda d)||^2 / 2
    (Dennis & Schnabel 1983, Algorithm A6.3.1): f0 = f(0), slope = f'(0) < 0, f_lam =
    f(lam) of the rejected step.  The first backtrack (or every one, order "quadratic")
    minimises the quadratic through f0, slope and f_lam; later ones the cubic through
    f0, slope, f(lam) and f(lam_prev).  The result is kept in [0.1 lam, 0.5 lam]."""
    if order == "quadratic" or lam_prev is None:
        den = 2.0 * (f_lam - f0 - slope * lam)
        lt = -slope * lam * lam / den if den > 0.0 else 0.5 * lam
    else:
        t1 = f_lam - f0 - lam * slope
        t2 = f_prev - f0 - lam_prev * slope
        a = (t1 / (lam * lam) - t2 / (lam_prev * lam_prev)) / (lam - lam_prev)
        b = (-lam_prev * t1 / (lam * lam) + lam * t2 / (lam_prev * lam_prev)) / (lam - lam_prev)
        if a == 0.0:
            lt = -slope / (2.0 * b) if b != 0.0 else 0.5 * lam
        else:
            d = max(b * b - 3.0 * a * slope, 0.0)
            lt = (-b + d ** 0.5) / (3.0 * a)
    if not np.isfinite(lt) or lt > 0.5 * lam:
        lt = 0.5 * lam
    return max(lt, 0.1 * lam)


@dataclass
class NewtonInfo:
    iterations: int = 0
    linear_iterations: int = 0
    fnorm0: float = 0.0
    fnorm: float = 0.0
    reason: str = ""
    history: list = field(default_factory=list)
    steps: list = field(default_factory=list)     # line-search step length of each Newton step
    energy: list = field(default_factory=list)    # step . F(new) per step (energy-norm test)


class MRNewton:
    """Newton iteration on the free DOFs of an MRSystem with a fixed set of
    constrained DOFs (their values are set in the state before each solve)."""

    def __init__(self, sys: MRSystem, fixed_mask: np.ndarray, material: np.ndarray, restart: int = 100):
        self.sys = sys
        dev = sys.device
        self.red = fem.DirichletReduction(sys.csr, fixed_mask)
        # ~200 entries per row and few rows: scheduled SELL shares each 64-row slice
        # between several wavefronts, which keeps every CU busy
        self.S = fem.SolverMatrix(self.red)
        self.restart = int(restart)
        self.ws = fem.KrylovWorkspace(max(self.red.n_free, 1), dev, restart=self.restart)
        self.free = self.red.free_idx.long()
        # c1, c2 per element (c1, c2, kappa for the nearly incompressible law)
        mat = np.ascontiguousarray(material, dtype=np.float64).reshape(
            -1, 3 if sys.formulation == NEARLY_INCOMPRESSIBLE else 2)
        self.material = torch.as_tensor(mat, device=dev)
        mu = 2.0 * (mat[:, 0] + mat[:, 1])
        pm = sys.pressure_mass(np.where(mu > 0, mu, 1.0))[~np.asarray(fixed_mask, dtype=bool)]
        pc = np.where(pm > 0, 1.0 / np.where(pm > 0, pm, 1.0), 0.0)
        self.pc_pos = torch.as_tensor(pc, device=dev)           # SPD form (MINRES)
        self.pc_neg = torch.as_tensor(-pc, device=dev)          # GMRES: negative Schur block
        self.R = torch.zeros(sys.ndofs, dtype=torch.float64, device=dev)
        self.f = torch.zeros(sys.ndofs, dtype=torch.float64, device=dev)
        self.J = sys.csr.zeros()
        self.Jff = None
        self.Jff32 = None

    def _eval(self, x, body_force, jac: bool, fd_step=None):
        self.R.zero_()
        self.f.zero_()
        if jac:
            self.J.zero_()
        self.sys.assemble(x, self.material, body_force=body_force, residual=self.R,
                          jac=self.J if jac and fd_step is None else None,
                          rhs=self.f if body_force is not None else None)
        if jac and fd_step is not None:
            self._jacobian_fd(x, body_force, fd_step)
        return (self.R - self.f).index_select(0, self.free)

    def _jacobian_fd(self, x, body_force, step):
        """Element Jacobians by one-sided finite differences, as
        EquationsSet_FiniteElementJacobianEvaluateFD (equations_set_routines.F90:2839-2962):
        column a of element e is (R_e(x + delta e_a) - R_e(x)) / delta with
        delta = (1 + ||x||_2) step, added into the equations CSR values.  Elements of one
        colour share no DOF, so column a of every element of a colour comes from one
        residual assembly of that colour with all their a-th DOFs perturbed."""
        sysm = self.sys
        cs = sysm.csr
        E, nd = cs.n_elem, cs.n_edof
        delta = (1.0 + float(torch.linalg.vector_norm(x).item())) * float(step)
        ed = cs.elem_dofs.long()
        pos = cs.pos.view(E, nd, nd)
        dev = x.device
        R0 = torch.zeros_like(self.R)
        f0 = torch.zeros_like(self.R)
        Ra = torch.zeros_like(self.R)
        fa = torch.zeros_like(self.R)
        bf = body_force
        for c in range(cs.n_colours):
            b0, b1 = int(cs.colour_ptr[c]), int(cs.colour_ptr[c + 1])
            if b1 == b0:
                continue
            els = cs.order[b0:b1].long()
            rows = ed[els]                                     # [ne_c, nd]
            cols = pos[els].permute(2, 0, 1).reshape(nd, -1)  # [a] -> positions of column a, rows in order
            R0.zero_()
            f0.zero_()
            sysm.assemble_colour(c, x, self.material, bf, residual=R0, rhs=f0 if bf is not None else None)
            base = R0 - f0
            dvec = torch.full((b1 - b0,), delta, dtype=torch.float64, device=dev)
            for a in range(nd):
                xp = x.clone()
                xp.index_add_(0, rows[:, a], dvec)
                Ra.zero_()
                fa.zero_()
                sysm.assemble_colour(c, xp, self.material, bf, residual=Ra, rhs=fa if bf is not None else None)
                diff = ((Ra - fa) - base) / delta
                self.J.index_add_(0, cols[a].long(), diff[rows].reshape(-1))

    @staticmethod
    def _ew_eta(fn, fn_prev, eta_prev, fn0, o):
        """Eisenstat-Walker forcing term, choice 2 with PETSc's SNES defaults
        (eta_0 = 0.3, gamma = 1, alpha = (1+sqrt 5)/2, safeguard threshold 0.1,
        eta_max = 0.9), floored at half the Newton target over ||F|| so the linear
        solve never goes below what the stopping test needs."""
        alpha = 0.5 * (1.0 + 5.0 ** 0.5)
        if fn_prev is None:
            eta = 0.3
        else:
            eta = (fn / fn_prev) ** alpha
            safe = eta_prev ** alpha
            if safe > 0.1:
                eta = max(eta, safe)
        eta = min(eta, 0.9)
        target = max(o.atol, o.rtol * fn0)
        return max(eta, 0.5 * target / max(fn, 1e-300), 1e-14)

    def _line_search(self, x, dx, F, fn, body_force, o):
        """Step length along the Newton step dx from x (F = F(x), fn = ||F||); returns
        (lambda, the new state, stalled)."""
        xt = x.clone()

        def trial(lam):
            xt.copy_(x)
            xt.index_add_(0, self.free, dx, alpha=lam)
            Ft = self._eval(xt, body_force, False)
            return Ft, float(torch.linalg.vector_norm(Ft).item())

        if o.linesearch == "basic":
            trial(1.0)
            return 1.0, xt, False
        if o.linesearch == "linear":
            # one secant step on s(lambda) = F(x + lambda dx) . dx, s(0) = F . dx
            s0 = float(torch.dot(F, dx).item())
            F1, _ = trial(1.0)
            s1 = float(torch.dot(F1, dx).item())
            lam = s0 / (s0 - s1) if s0 != s1 else 1.0
            if not np.isfinite(lam) or lam <= o.ls_minlambda:
                lam = 1.0
            _, ft = trial(lam)
            return lam, xt, not np.isfinite(ft)
        f0 = 0.5 * fn * fn
        slope = -fn * fn          # f'(0) along an exact Newton step
        lam, lam_prev, f_prev = 1.0, None, None
        for _ in range(max(o.ls_maxit, 1)):
            _, ft = trial(lam)
            f_lam = 0.5 * ft * ft if np.isfinite(ft) else np.inf
            if f_lam <= f0 + o.ls_alpha * lam * slope:
                return lam, xt, False
            if not np.isfinite(f_lam):
                lam_new = 0.1 * lam
            else:
                lam_new = backtrack_lambda(o.linesearch, lam, f0, slope, f_lam, lam_prev, f_prev)
            lam_prev, f_prev, lam = lam, f_lam, lam_new
            if lam < o.ls_minlambda:
                break
        return lam, xt, True

    def internal_residual(self, x):
        self.R.zero_()
        self.sys.assemble(x, self.material, residual=self.R)
        return self.R.clone()

    def solve(self, x, body_force=None, opts: NewtonOptions | None = None) -> NewtonInfo:
        """Newton with backtracking on ||F||, F = R_int(x) - f_gravity(x) on the free rows;
        x (device, FP64, all DOFs) is updated in place."""
        o = opts or NewtonOptions()
        info = NewtonInfo()
        if self.red.n_free == 0:
            info.reason = "no free DOFs"
            return info
        linear = o.linear
        if linear == "auto":
            # the deviatoric tangent is symmetric; the general branch's is not, nor is the
            # nearly incompressible one when c2 != 0 (its c2 term carries J^(-2/3) where a
            # strain energy would give J^(-4/3))
            linear = "minres" if self.sys.formulation == DEVIATORIC else "gmres"
        fd = o.fd_step if o.jacobian == "fd" else None
        if o.convergence not in ("petsc", "energy"):
            # SOLVER_NEWTON_CONVERGENCE_DIFFERENTIATED_RATIO is not implemented in iron either
            raise CMFEError(1, f"convergence test {o.convergence!r} is not implemented")
        energy_done, energy0 = False, 1.0
        F = self._eval(x, body_force, True, fd)
        fn = float(torch.linalg.vector_norm(F).item())
        info.fnorm0 = fn
        info.history.append(fn)
        ew = o.linear_rtol == "ew"
        eta_prev, fn_prev = 0.3, None
        while True:
            if not np.isfinite(fn):
                raise CMFEError(1, "Nonlinear line search solver did not converge. PETSc diverged F Norm NaN.")
            if o.convergence == "energy":
                # F = 0 exactly: the step is zero, and so is its energy
                if energy_done or fn == 0.0:
                    info.reason = "converged energy norm"
                    break
            elif fn <= o.atol:
                info.reason = "converged F norm absolute"
                break
            elif info.iterations > 0 and fn <= o.rtol * info.fnorm0:
                info.reason = "converged F norm relative"
                break
            if info.iterations >= o.maxit:
                raise CMFEError(1, "Nonlinear line search solver did not converge. PETSc diverged maximum "
                                   "iterations.")
            dx = torch.zeros(self.red.n_free, dtype=torch.float64, device=x.device)
            b = -F
            if ew:
                eta = self._ew_eta(fn, fn_prev, eta_prev, info.fnorm0, o)
                eta_prev, fn_prev = eta, fn
            else:
                eta = float(o.linear_rtol)
            if linear == "minres" and o.minres_values == "f32":
                self.Jff32 = self.S.gather(self.J, fem.F32, out=self.Jff32)
                vals = self.Jff32
            else:
                self.Jff = self.S.gather(self.J, fem.F64, out=self.Jff)
                vals = self.Jff
            if linear == "minres":
                li = fem.minres(self.S, vals, b, dx, eta, o.linear_maxit, self.ws, pc_inv=self.pc_pos)
            else:
                li = fem.gmres(self.S, vals, b, dx, eta, o.linear_maxit, self.ws,
                               restart=o.gmres_restart, pc_inv=self.pc_neg)
            info.linear_iterations += li.iterations
            lam, xt, stalled = self._line_search(x, dx, F, fn, body_force, o)
            if stalled:
                # no decrease along an exact Newton direction: ||F|| sits at round-off
                if fn <= 1e-10 * max(info.fnorm0, 1.0):
                    info.reason = "converged F norm at round-off"
                    break
                raise CMFEError(1, "Nonlinear line search solver did not converge. PETSc diverged line search.")
            snorm = lam * float(torch.linalg.vector_norm(dx).item())
            x.copy_(xt)
            info.iterations += 1
            info.steps.append(lam)
            F = self._eval(x, body_force, True, fd)
            fn = float(torch.linalg.vector_norm(F).item())
            info.history.append(fn)
            if o.convergence == "energy":
                # energy = step . F(new): relative to its first value, below atol; a first
                # value at zero (5 eps, iron's ZERO_TOLERANCE) converges at once
                energy = -lam * float(torch.dot(dx, F).item())
                if info.iterations == 1:
                    energy0 = energy
                    energy_done = abs(energy) < 5.0 * np.finfo(np.float64).eps
                else:
                    energy_done = abs(energy / energy0) < o.atol
                info.energy.append(energy)
                continue
            xnorm = float(torch.linalg.vector_norm(x.index_select(0, self.free)).item())
            if snorm < o.stol * xnorm and fn > o.atol and fn > o.rtol * info.fnorm0:
                info.reason = "converged S norm relative"
                break
        info.fnorm = fn
        return info


def load_increments(newton: MRNewton, x, n_increments: int, incremented: np.ndarray | None = None,
                    full_value: np.ndarray | None = None, body_force=None, opts: NewtonOptions | None = None):
    """The load-increment control loop: iteration i of N moves each FIXED_INCREMENTED
    DOF to current + (full - current)/(N - i + 1) (equations_set_routines.F90:6340-6350)
    and scales the gravity body force by i/N when N > 1
    (finite_elasticity_routines.F90:13333-13350), then runs Newton."""
    infos = []
    dev = x.device
    inc_idx = None
    if incremented is not None and np.any(incremented):
        inc_idx = torch.as_tensor(np.nonzero(incremented)[0], device=dev)
        full = torch.as_tensor(np.asarray(full_value, dtype=np.float64)[incremented], device=dev)
    for i in range(1, n_increments + 1):
        if inc_idx is not None:
            cur = x.index_select(0, inc_idx)
            x.index_copy_(0, inc_idx, cur + (full - cur) / float(n_increments - i + 1))
        bf = None
        if body_force is not None:
            scale = i / n_increments if n_increments > 1 else 1.0
            bf = [float(v) * scale for v in body_force]
        infos.append(newton.solve(x, bf, opts))
    return infos


# ------------------------------------------------------- iron API solve
def solve(problem, es, solver, bc):
    """Problem.Solve() for a FINITE_ELASTICITY equations set (MOONEY_RIVLIN subtype,
    NO_SUBTYPE for the general branch, or NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN with a
    3-component dependent field and materials c1, c2, kappa)."""
    from . import engine
    from . import iron
    rm: RegularMesh = es.geometricField._mesh()
    dep = es.dependentField
    U = iron.FieldVariableTypes.U
    var = dep._var(U)
    sub = es.specification[2]
    nearly = sub == iron.EquationsSetSubtypes.NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN
    if nearly and var.ncomp != 3:
        raise CMFEError(1, "the nearly incompressible dependent field has 3 displacement components")
    if not nearly and var.ncomp != 4:
        raise CMFEError(1, "the finite-elasticity dependent field needs 3 displacement components and a pressure")
    if any(mc != 1 for mc in var.mesh_comp[:3]) or any(
            it != iron.FieldInterpolationTypes.NODE_BASED for it in var.interp[:3]):
        raise CMFEError(1, "displacement components must be node based on mesh component 1")
    if nearly:
        p_mc = None
    elif var.interp[3] == iron.FieldInterpolationTypes.ELEMENT_BASED:
        p_mc = None
    elif var.interp[3] == iron.FieldInterpolationTypes.NODE_BASED:
        p_mc = var.mesh_comp[3] - 1
    else:
        raise CMFEError(1, "the pressure must be element based or node based")
    if nearly:
        form = NEARLY_INCOMPRESSIBLE
    else:
        form = DEVIATORIC if sub == iron.EquationsSetSubtypes.MOONEY_RIVLIN else CAUCHY
    # conditions on any other field are an error, as on the linear path
    if any(not engine._is(entry[0], dep) for entry in bc._entries):
        raise CMFEError(1, "boundary conditions must be set on the dependent field")
    coords = engine.geometry_coords(es.geometricField, rm, 0)
    cache = getattr(es, "_mr_cache", None)
    key = (id(rm), p_mc, form, coords.tobytes())
    if cache is None or cache[0] != key:
        sys = MRSystem(rm, p_mc, form, coords=coords)
        cache = (key, sys, {})
        es._mr_cache = cache
    sys = cache[1]
    # boundary conditions on U: FIXED sets/adds the value, FIXED_INCREMENTED ramps to it
    n = sys.ndofs
    if dep._data(U).shape[0] != n:
        raise CMFEError(1, "dependent field layout does not match the finite-elasticity DOF layout")
    x_host = dep._data(U).copy()
    fixed = np.zeros(n, dtype=bool)
    incremented = np.zeros(n, dtype=bool)
    full = np.zeros(n)
    if getattr(bc, "_constraints", None):
        raise CMFEError(1, "DOF constraints are implemented for linear equations sets only")
    for (fld, vt, node, comp, cond, value, add) in bc._entries:
        if vt != U:
            raise CMFEError(1, "finite elasticity supports boundary conditions on the U variable only")
        if comp > 3 and p_mc is None:
            raise CMFEError(1, "element-based pressure DOFs cannot carry node boundary conditions")
        mc = var.mesh_comp[comp - 1] - 1
        dof = int(dep._offset(U, comp) + rm.local_node(node, mc))
        fixed[dof] = True
        if cond == iron.BoundaryConditionsTypes.FIXED:
            x_host[dof] = (x_host[dof] + value) if add else value
        elif cond == iron.BoundaryConditionsTypes.FIXED_INCREMENTED:
            incremented[dof] = True
            full[dof] = (dep._data(U)[dof] + value) if add else value
        else:
            raise CMFEError(1, f"boundary condition type {cond} is not supported for finite elasticity")
    # materials: c1, c2 (and kappa) per element; density (V) and gravity (source) -> body force
    mf = es.materialsField
    mat = np.stack([engine._element_values(mf, c, rm) for c in ((1, 2, 3) if nearly else (1, 2))], axis=1)
    body_force = None
    if es.sourceField is not None and mf is not None and iron.FieldVariableTypes.V in mf._vars:
        rho = engine._element_values(mf, 1, rm, vt=iron.FieldVariableTypes.V)
        g = np.stack([engine._element_values(es.sourceField, c, rm) for c in (1, 2, 3)], axis=1)
        bf = rho[:, None] * g
        if not np.allclose(bf, bf[0]):
            raise CMFEError(1, "a spatially varying body force is not supported")
        if np.any(bf[0] != 0):
            body_force = bf[0]
    fkey = (fixed.tobytes(), mat.tobytes())
    newton = cache[2].get(fkey)
    if newton is None:
        newton = MRNewton(sys, fixed, mat)
        cache[2].clear()
        cache[2][fkey] = newton
    lin = solver._linear
    opts = NewtonOptions(atol=float(solver.newtonAbsoluteTolerance), rtol=float(solver.newtonRelativeTolerance),
                         stol=float(getattr(solver, "newtonSolutionTolerance", 1e-5)),
                         maxit=int(solver.newtonMaximumIterations))
    if lin is not None:
        if int(lin.linearType) == iron.LinearSolverTypes.DIRECT:
            # no GPU factorisation: iterate to round-off with a long Krylov basis instead
            opts.linear_rtol = 1e-12
            opts.gmres_restart = 100
        else:
            opts.linear_rtol = float(lin.linearIterativeRelativeTolerance)
            opts.linear_maxit = int(lin.linearIterativeMaximumIterations)
            opts.gmres_restart = int(lin.linearIterativeGMRESRestart)
    ls_type = int(getattr(solver, "newtonLineSearchType", iron.NewtonLineSearchTypes.CUBIC))
    opts.linesearch = {iron.NewtonLineSearchTypes.NONORMS: "basic", iron.NewtonLineSearchTypes.LINEAR: "linear",
                       iron.NewtonLineSearchTypes.QUADRATIC: "quadratic",
                       iron.NewtonLineSearchTypes.CUBIC: "cubic"}[iron.NewtonLineSearchTypes(ls_type)]
    opts.ls_alpha = float(getattr(solver, "newtonLineSearchAlpha", 1e-4))
    ct = int(getattr(solver, "newtonConvergenceTestType", iron.NewtonConvergenceTypes.PETSC_DEFAULT))
    opts.convergence = {iron.NewtonConvergenceTypes.PETSC_DEFAULT: "petsc",
                        iron.NewtonConvergenceTypes.ENERGY_NORM: "energy",
                        iron.NewtonConvergenceTypes.DIFFERENTIATED_RATIO: "differentiated_ratio"}[
        iron.NewtonConvergenceTypes(ct)]
    if int(getattr(solver, "newtonJacobianCalculationType", iron.JacobianCalculationTypes.EQUATIONS)) == \
            iron.JacobianCalculationTypes.FD:
        opts.jacobian = "fd"
        eq = es.equations
        opts.fd_step = float(getattr(eq, "jacobianFiniteDifferenceStepSize", 1e-6)) if eq is not None else 1e-6
    x = torch.as_tensor(x_host, device=sys.device)
    n_inc = max(int(problem._control_loop.maximumIterations), 1)
    infos = load_increments(newton, x, n_inc, incremented, full, body_force, opts)
    # EQUATIONS_SET_NONLINEAR_RHS_UPDATE: DELUDELN = assembled residual (RHS-free DOFs)
    R = newton.internal_residual(x)
    dep._data(U)[:] = x.cpu().numpy()
    dep._data(iron.FieldVariableTypes.DELUDELN)[:] = R.cpu().numpy()
    solver.last_info = infos
    return infos
