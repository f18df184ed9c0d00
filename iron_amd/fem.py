"""Device-resident finite-element linear systems on top of libmi_fe.so.

This is the engine behind the iron-compatible API (iron_amd.iron): it holds the
mesh, the CSR structure, the scatter map and the Dirichlet reduction on the GPU
and drives the HIP kernels through the C ABI.  PyTorch only provides device
memory and the stream.

Pipeline of a static linear solve (reference call stack in parentheses):
  assemble()      element integrals + coloured scatter-add into the equations CSR
                  (EquationsSet_AssembleStaticLinearFEM -> ElementAdd)
  reduce.gather() K_ff copy into the solver matrix (SOLVER_MATRIX_EQUATIONS_MATRIX_ADD)
  reduce.rhs()    b_f = s_f + f_f - K_fd u_d (SOLVER_MATRICES_STATIC_ASSEMBLE RHS)
  solve_*()       Jacobi CG / mixed refinement / GMRES (Petsc_KSPSolve)
  field_update()  u_f = x (SOLVER_VARIABLES_FIELD_UPDATE)
  backsub()       reactions on the RHS variable (EQUATIONS_SET_BACKSUBSTITUTE)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import ElementDesc, MatrixDesc, call, ptr
from .basis import LagrangeBasis

F64, F32 = _lib.MIFE_F64, _lib.MIFE_F32


def _dev(device=None):
    if not torch.cuda.is_available():
        raise RuntimeError("iron_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device(device if device is not None else "cuda:0")


def _stream():
    return _lib.stream_ptr()


class DeviceBasis:
    """Basis tables uploaded once: dN [ng][nn][dim], N [ng][nn], w [ng]."""

    def __init__(self, b: LagrangeBasis, device):
        self.basis = b
        self.dN = torch.as_tensor(np.ascontiguousarray(b.dN), dtype=torch.float64, device=device)
        self.N = torch.as_tensor(np.ascontiguousarray(b.N), dtype=torch.float64, device=device)
        self.w = torch.as_tensor(np.ascontiguousarray(b.weights), dtype=torch.float64, device=device)
        self.desc = ElementDesc(b.num_xi, b.num_nodes, b.num_gauss, ptr(self.dN), ptr(self.N), ptr(self.w))


class CSRStructure:
    """FEM sparsity of an equations matrix plus the per-element scatter map and
    the element colouring used for race-free, deterministic scatter-adds."""

    def __init__(self, elem_dofs: np.ndarray, n_dofs: int, device=None, colours: np.ndarray | None = None):
        dev = _dev(device)
        self.device = dev
        self.n = int(n_dofs)
        ed = np.ascontiguousarray(elem_dofs, dtype=np.int32)
        self.n_elem, self.n_edof = ed.shape
        self.elem_dofs = torch.as_tensor(ed, device=dev)
        s = _stream()
        L = _lib.lib()
        wbytes = L.mife_csr_pattern_work_bytes(self.n_elem, self.n_edof, self.n)
        work = torch.empty(int(wbytes), dtype=torch.uint8, device=dev)
        self.row_ptr = torch.empty(self.n + 1, dtype=torch.int32, device=dev)
        nnz = C.c_int64(0)
        call("mife_csr_pattern", ptr(self.elem_dofs), self.n_elem, self.n_edof, self.n, ptr(self.row_ptr), None,
             C.byref(nnz), ptr(work), s)
        self.nnz = int(nnz.value)
        self.col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        call("mife_csr_pattern", ptr(self.elem_dofs), self.n_elem, self.n_edof, self.n, ptr(self.row_ptr),
             ptr(self.col), None, ptr(work), s)
        del work
        self._pos = None
        if colours is None:
            col_h = np.empty(self.n_elem, dtype=np.int32)
            nc = C.c_int(0)
            call("mife_colour_elements", ed.ctypes.data, self.n_elem, self.n_edof, self.n, col_h.ctypes.data,
                 C.byref(nc))
            colours = col_h
        colours = np.asarray(colours, dtype=np.int32)
        order = np.argsort(colours, kind="stable").astype(np.int32)
        counts = np.bincount(colours, minlength=int(colours.max()) + 1)
        self.colour_ptr = np.zeros(len(counts) + 1, dtype=np.int32)
        self.colour_ptr[1:] = np.cumsum(counts)
        self.n_colours = len(counts)
        self.order = torch.as_tensor(order, device=dev)

    @property
    def pos(self) -> torch.Tensor:
        """Per-element CSR position map [E][n_edof][n_edof] (built on first use: the
        structured grid assembly does without it)."""
        if self._pos is None:
            self._pos = torch.empty(self.n_elem * self.n_edof * self.n_edof, dtype=torch.int32, device=self.device)
            call("mife_scatter_map", ptr(self.elem_dofs), self.n_elem, self.n_edof, ptr(self.row_ptr), ptr(self.col),
                 ptr(self._pos), _stream())
        return self._pos

    def zeros(self, prec=F64):
        return torch.zeros(max(self.nnz, 1), dtype=torch.float64 if prec == F64 else torch.float32, device=self.device)

    def to_host(self):
        """(row_ptr, col) as int64 numpy arrays."""
        return self.row_ptr.cpu().numpy().astype(np.int64), self.col[: self.nnz].cpu().numpy().astype(np.int64)


class DirichletReduction:
    """Free-DOF solver system: rows/columns of free DOFs in ascending order."""

    def __init__(self, csr: CSRStructure, fixed_mask: np.ndarray):
        dev = csr.device
        self.csr = csr
        fixed = np.asarray(fixed_mask, dtype=bool)
        free = np.nonzero(~fixed)[0].astype(np.int32)
        new_idx = np.full(csr.n, -1, dtype=np.int32)
        new_idx[free] = np.arange(len(free), dtype=np.int32)
        self.n_free = len(free)
        self.fixed_mask = fixed
        self.free_idx = torch.as_tensor(free, device=dev)
        self.new_idx = torch.as_tensor(new_idx, device=dev)
        s = _stream()
        wbytes = _lib.lib().mife_csr_pattern_work_bytes(1, 1, csr.n)
        work = torch.empty(int(wbytes), dtype=torch.uint8, device=dev)
        self.row_ptr = torch.empty(self.n_free + 1, dtype=torch.int32, device=dev)
        nnz = C.c_int64(0)
        call("mife_reduce_pattern", csr.n, ptr(csr.row_ptr), ptr(csr.col), ptr(self.new_idx), self.n_free,
             ptr(self.row_ptr), None, None, C.byref(nnz), ptr(work), s)
        self.nnz = int(nnz.value)
        self.col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        self.src_pos = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        call("mife_reduce_pattern", csr.n, ptr(csr.row_ptr), ptr(csr.col), ptr(self.new_idx), self.n_free,
             ptr(self.row_ptr), ptr(self.col), ptr(self.src_pos), None, ptr(work), s)
        # rows with no fixed column have no lifting term: flagged once here, the RHS
        # kernel skips their rows (the same sums, without reading them)
        self.no_lift = torch.ones(csr.n, dtype=torch.uint8, device=dev)
        if csr.nnz and self.n_free < csr.n:
            fixed_col = torch.index_select(self.new_idx, 0, csr.col[: csr.nnz]) < 0
            cum = torch.cumsum(fixed_col, 0, dtype=torch.int32)
            cum = torch.cat([torch.zeros(1, dtype=torch.int32, device=dev), cum])
            rp = csr.row_ptr.long()
            self.no_lift = ((cum[rp[1:]] - cum[rp[:-1]]) == 0).to(torch.uint8)
            del fixed_col, cum

    def gather(self, vals: torch.Tensor, prec=F64, out: torch.Tensor | None = None) -> torch.Tensor:
        """K_ff values from equations-CSR values (FP64 or FP32) in precision ``prec``."""
        if out is None:
            out = torch.empty(max(self.nnz, 1), dtype=torch.float64 if prec == F64 else torch.float32,
                              device=vals.device)
        src_prec = F64 if vals.dtype == torch.float64 else F32
        call("mife_gather_values", self.nnz, ptr(self.src_pos), ptr(vals), src_prec, ptr(out), prec, _stream())
        return out

    def rhs(self, vals64, u, src=None, rhs=None, rhs_fixed=None, out=None) -> torch.Tensor:
        """b_f = src + rhs - K_fd u_d (no lifting on rows flagged in ``rhs_fixed``, uint8)."""
        if out is None:
            out = torch.empty(max(self.n_free, 1), dtype=torch.float64, device=vals64.device)
        skip = self.no_lift if rhs_fixed is None else (self.no_lift | rhs_fixed)
        call("mife_solver_rhs", self.n_free, ptr(self.free_idx), ptr(self.csr.row_ptr), ptr(self.csr.col),
             ptr(vals64), ptr(self.new_idx), ptr(u), ptr(src), ptr(rhs), ptr(skip), ptr(out), _stream())
        return out

    def field_update(self, x, u):
        call("mife_field_update", self.n_free, ptr(self.free_idx), ptr(x), ptr(u), _stream())


class CoupledReduction:
    """Solver system with linear DOF constraints, P^T K P.

    ``owner[g]`` is the solver DOF of equations DOF g (-1 when Dirichlet-fixed) and
    ``coef[g]`` its coupling coefficient: a constrained DOF u_c = a u_m takes m's
    solver DOF and coefficient a, every other free DOF coefficient 1
    (SOLVER_MAPPING_CALCULATE with DOF couplings, solver_mapping_routines.F90:445-520,
    :1445-1511).  Solver values are sums of coupled equations entries
    (``mife_couple_values``), so ``src_pos`` indexes this object's own solver-CSR
    values: a SELL solver matrix gathers from them.
    """

    coupled = True

    def __init__(self, csr: CSRStructure, owner: np.ndarray, coef: np.ndarray, constrained: np.ndarray):
        dev = csr.device
        self.csr = csr
        owner = np.asarray(owner, dtype=np.int32)
        constrained = np.asarray(constrained, dtype=bool)
        if owner.shape != (csr.n,) or np.shape(coef) != (csr.n,) or constrained.shape != (csr.n,):
            raise ValueError("owner, coef and constrained need one entry per equations DOF")
        self.n_free = int(owner.max()) + 1 if csr.n and owner.max() >= 0 else 0
        self.fixed_mask = owner < 0
        # the solver DOFs are the unconstrained free DOFs in ascending order
        lead = np.nonzero((owner >= 0) & ~constrained)[0].astype(np.int32)
        if len(lead) != self.n_free or not np.array_equal(owner[lead], np.arange(self.n_free)):
            raise ValueError("owner must number the unconstrained free DOFs 0, 1, ... in ascending order")
        self.lead = lead
        self.owner = torch.as_tensor(owner, device=dev)
        self.coef = torch.as_tensor(np.asarray(coef, dtype=np.float64), device=dev)
        self.free_idx = torch.as_tensor(lead, device=dev)
        n_s = self.n_free
        self.row_ptr = torch.zeros(n_s + 1, dtype=torch.int32, device=dev)
        self.grp_ptr = torch.zeros(n_s + 1, dtype=torch.int32, device=dev)
        n_grp = int((owner >= 0).sum())
        self.grp_rows = torch.empty(max(n_grp, 1), dtype=torch.int32, device=dev)
        self.nnz, self.n_contrib = 0, 0
        self.col = torch.zeros(1, dtype=torch.int32, device=dev)
        self.contrib_ptr = torch.zeros(1, dtype=torch.int32, device=dev)
        self.contrib_pos = self.contrib_row = torch.zeros(1, dtype=torch.int32, device=dev)
        if n_s > 0 and csr.nnz > 0:
            s = _stream()
            wbytes = _lib.lib().mife_couple_pattern_work_bytes(csr.n, csr.nnz, n_s)
            work = torch.empty(int(wbytes), dtype=torch.uint8, device=dev)
            nnz_s, n_c = C.c_int64(0), C.c_int64(0)
            call("mife_couple_pattern", csr.n, csr.nnz, ptr(csr.row_ptr), ptr(csr.col), ptr(self.owner), n_s,
                 ptr(self.grp_ptr), ptr(self.grp_rows), ptr(self.row_ptr), None, None, None, None, C.byref(nnz_s),
                 C.byref(n_c), ptr(work), s)
            self.nnz, self.n_contrib = int(nnz_s.value), int(n_c.value)
            self.col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
            self.contrib_ptr = torch.empty(self.nnz + 1, dtype=torch.int32, device=dev)
            self.contrib_pos = torch.empty(max(self.n_contrib, 1), dtype=torch.int32, device=dev)
            self.contrib_row = torch.empty(max(self.n_contrib, 1), dtype=torch.int32, device=dev)
            call("mife_couple_pattern", csr.n, csr.nnz, ptr(csr.row_ptr), ptr(csr.col), ptr(self.owner), n_s,
                 ptr(self.grp_ptr), ptr(self.grp_rows), ptr(self.row_ptr), ptr(self.col), ptr(self.contrib_ptr),
                 ptr(self.contrib_pos), ptr(self.contrib_row), None, None, ptr(work), s)
            del work
        self.src_pos = torch.arange(max(self.nnz, 1), dtype=torch.int32, device=dev)

    def gather(self, vals: torch.Tensor, prec=F64, out: torch.Tensor | None = None) -> torch.Tensor:
        """Solver-CSR values of P^T K P (FP64 sums) from equations-CSR values."""
        if out is None:
            out = torch.empty(max(self.nnz, 1), dtype=torch.float64 if prec == F64 else torch.float32,
                              device=vals.device)
        src_prec = F64 if vals.dtype == torch.float64 else F32
        call("mife_couple_values", self.nnz, ptr(self.contrib_ptr), ptr(self.contrib_pos), ptr(self.contrib_row),
             ptr(self.csr.col), ptr(self.coef), ptr(vals), src_prec, ptr(out), prec, _stream())
        return out

    def rhs(self, vals64, u, src=None, rhs=None, rhs_fixed=None, out=None) -> torch.Tensor:
        """b_s = sum over the rows coupled to s of a_r (src_r + rhs_r - lift_r)."""
        if out is None:
            out = torch.empty(max(self.n_free, 1), dtype=torch.float64, device=vals64.device)
        call("mife_solver_rhs_coupled", self.n_free, ptr(self.grp_ptr), ptr(self.grp_rows), ptr(self.coef),
             ptr(self.csr.row_ptr), ptr(self.csr.col), ptr(vals64), ptr(self.owner), ptr(u), ptr(src), ptr(rhs),
             ptr(rhs_fixed), ptr(out), _stream())
        return out

    def field_update(self, x, u):
        call("mife_field_update_coupled", self.csr.n, ptr(self.owner), ptr(self.coef), ptr(x), ptr(u), _stream())


class BlockReduction:
    """Several equations sets in one solver system.

    On one rank SOLVER_MAPPING_CALCULATE numbers the solver rows set by set, each
    set's included rows in ascending order (solver_mapping_routines.F90:791-830), and
    the columns likewise; with no interface conditions the solver matrix is block
    diagonal.  ``parts`` are the per-set reductions (Dirichlet or coupled); values,
    RHS and field updates are taken per part into slices of the joint vectors.
    """

    coupled = True

    def __init__(self, parts):
        self.parts = list(parts)
        dev = self.parts[0].csr.device
        self.row_off = np.cumsum([0] + [p.n_free for p in self.parts]).tolist()
        self.nnz_off = np.cumsum([0] + [p.nnz for p in self.parts]).tolist()
        self.n_free = self.row_off[-1]
        self.nnz = self.nnz_off[-1]
        rps, cols = [], []
        for p, r0, z0 in zip(self.parts, self.row_off, self.nnz_off):
            rps.append(p.row_ptr[:-1].to(torch.int64) + z0)
            cols.append(p.col[: p.nnz].to(torch.int64) + r0)
        rps.append(torch.tensor([self.nnz], dtype=torch.int64, device=dev))
        self.row_ptr = torch.cat(rps).to(torch.int32)
        self.col = torch.cat(cols).to(torch.int32) if self.nnz else torch.zeros(1, dtype=torch.int32, device=dev)
        self.src_pos = torch.arange(max(self.nnz, 1), dtype=torch.int32, device=dev)
        self.free_idx = None

    def gather(self, vals, prec=F64, out: torch.Tensor | None = None) -> torch.Tensor:
        """Joint solver-CSR values from the per-set equations values ``vals`` (a list)."""
        if out is None:
            out = torch.empty(max(self.nnz, 1), dtype=torch.float64 if prec == F64 else torch.float32,
                              device=vals[0].device)
        for p, v, z0 in zip(self.parts, vals, self.nnz_off):
            if p.nnz:
                p.gather(v, prec, out=out[z0:z0 + p.nnz])
        return out

    def rhs(self, parts_args, out=None) -> torch.Tensor:
        """Joint RHS: ``parts_args`` holds (vals64, u, src, rhs, rhs_fixed) per set."""
        dev = self.parts[0].csr.device
        if out is None:
            out = torch.empty(max(self.n_free, 1), dtype=torch.float64, device=dev)
        for p, a, r0 in zip(self.parts, parts_args, self.row_off):
            if p.n_free:
                p.rhs(a[0], a[1], src=a[2], rhs=a[3], rhs_fixed=a[4], out=out[r0:r0 + p.n_free])
        return out

    def field_update(self, x, us):
        for p, u, r0 in zip(self.parts, us, self.row_off):
            if p.n_free:
                p.field_update(x[r0:r0 + p.n_free], u)


class SolverMatrix:
    """Storage of the reduced (free x free) solver matrix.

    ``fmt="csr"`` keeps iron's solver pattern as CSR; ``fmt="sell"`` (default)
    stores it in slices of 64 rows, one row per wavefront lane, rows sorted by
    length inside windows of ``sigma`` rows so a slice is nearly rectangular.
    A window of 1024 rows leaves 3% padding on the 48^3 triquadratic box (256
    rows: 12%, 64 rows: 46%) and widens a slice's column span by no more than the
    window, so 16-bit columns still fit.  Either way the values are gathered straight from the equations-matrix CSR.
    ``schedule`` gives SELL products a slice order: eight contiguous ranges of
    slices (one per XCD), each walked longest slice first.  ``interleave =
    (ncomp, n_per_comp)`` orders the rows of a vector field (iron's separated DOF
    order: component-major) node by node before the length sort, so that the
    columns of a slice stay within a narrow range (16-bit columns).  With
    ``block=True`` and ``interleave = (3, n_per_comp)`` a field whose free DOFs are
    whole nodes, the three rows of each node alike, is stored as SELL3: SELL over
    node rows with 3x3 blocks (one 16-bit node column per nine values, three
    contiguous x reads per block); other fields keep the scalar layout.
    """

    def __init__(self, red: DirichletReduction, fmt: str = "sell", sigma: int = 1024, compress: bool = True,
                 schedule: bool = True, interleave: tuple | None = None, block: bool = False):
        dev = red.row_ptr.device
        self.red = red
        self.n = red.n_free
        self.fmt = fmt
        self.block = 1  # rows per stored row of the layout (SELL3: 3)
        if fmt == "csr":
            self.src = red.src_pos
            self.stored = red.nnz
            self.padding = 0.0
            self.desc = MatrixDesc(self.n, _lib.MIFE_FMT_CSR, ptr(red.row_ptr), ptr(red.col), None, 0, red.nnz, None)
            return
        if fmt != "sell":
            raise ValueError(f"unknown solver matrix format {fmt!r}")
        if (block and compress and interleave is not None and int(interleave[0]) == 3 and self.n > 0
                and not getattr(red, "coupled", False) and self._build_sell3(red, int(interleave[1]), sigma,
                                                                             schedule)):
            return
        n = self.n
        if n == 0:  # every DOF fixed: an empty matrix (valid, non-null pointers)
            z = torch.zeros(1, dtype=torch.int32, device=dev)
            self.perm = self.iperm = torch.zeros(0, dtype=torch.int32, device=dev)
            self.slice_ptr = self.col = self.src = z
            self.col16 = self.col_base = self.slice_order = None
            self.stored, self.padding = 0, 0.0
            self.desc = MatrixDesc(0, _lib.MIFE_FMT_SELL, ptr(z), ptr(z), ptr(z), 0, 0, None, None, 0)
            return
        # slice order on the device: rows by window of sigma positions, then by
        # decreasing length, then by position (mife_sell_permutation)
        lengths = (red.row_ptr[1:] - red.row_ptr[:-1]).to(torch.int64)
        row_at = None
        if interleave is not None:
            # position of a row: rank of its (node, component) key among the free
            # DOFs' keys (keys are distinct), by a prefix sum over all keys
            ncomp, nper = int(interleave[0]), int(interleave[1])
            dof = red.free_idx.to(torch.int64)
            key = (dof % nper) * ncomp + dof // nper
            mark = torch.zeros(ncomp * nper, dtype=torch.int32, device=dev)
            mark[key] = 1
            pos = (torch.cumsum(mark, 0, dtype=torch.int32) - 1)[key]
            row_at = torch.empty(n, dtype=torch.int32, device=dev)
            row_at[pos.long()] = torch.arange(n, dtype=torch.int32, device=dev)
        self.perm = torch.empty(n, dtype=torch.int32, device=dev)
        call("mife_sell_permutation", n, ptr(red.row_ptr), ptr(row_at), int(sigma), ptr(self.perm), _stream())
        perm = self.perm.long()
        self.iperm = torch.empty_like(self.perm)
        self.iperm[perm] = torch.arange(n, dtype=torch.int32, device=dev)
        n_slices = (n + 63) // 64
        lp = torch.zeros(n_slices * 64, dtype=torch.int64, device=dev)
        lp[:n] = lengths[perm]
        slice_ptr = torch.zeros(n_slices + 1, dtype=torch.int64, device=dev)
        slice_ptr[1:] = torch.cumsum(lp.view(n_slices, 64).amax(dim=1) * 64, 0)
        self.stored = int(slice_ptr[-1].item())
        if self.stored >= 2 ** 31:
            raise ValueError("SELL storage exceeds the int32 index range")
        self.padding = self.stored / max(red.nnz, 1) - 1.0
        self.slice_ptr = slice_ptr.to(torch.int32)
        self.col = torch.empty(max(self.stored, 1), dtype=torch.int32, device=dev)
        self.src = torch.empty(max(self.stored, 1), dtype=torch.int32, device=dev)
        call("mife_sell_build", n, ptr(red.row_ptr), ptr(red.col), ptr(red.src_pos), ptr(self.perm), ptr(self.iperm),
             n_slices, ptr(self.slice_ptr), ptr(self.col), ptr(self.src), _stream())
        self.col16 = None
        self.col_base = None
        colp = ptr(self.col)
        if compress:
            # 16-bit slice-relative columns when every slice spans < 65,536 columns
            col16 = torch.empty(max(self.stored, 1), dtype=torch.int16, device=dev)
            cbase = torch.empty(max(n_slices, 1), dtype=torch.int32, device=dev)
            ok = C.c_int(0)
            call("mife_sell_compress", n_slices, ptr(self.slice_ptr), ptr(self.col), ptr(cbase), ptr(col16),
                 C.byref(ok), _stream())
            if ok.value:
                self.col16, self.col_base = col16, cbase
                colp = ptr(col16)
        self.slice_order = None
        span = C.c_int(0)
        if schedule and n_slices > 0:
            self.slice_order = torch.empty(_lib.lib().mife_sell_order_len(n_slices), dtype=torch.int32, device=dev)
            call("mife_sell_order", n_slices, ptr(self.slice_ptr), ptr(self.col_base), ptr(self.slice_order),
                 C.byref(span), _stream())
        self.desc = MatrixDesc(n, _lib.MIFE_FMT_SELL, ptr(self.slice_ptr), colp, ptr(self.perm), n_slices,
                               self.stored, ptr(self.col_base), ptr(self.slice_order), span.value)

    def _build_sell3(self, red, nper: int, sigma: int, schedule: bool) -> bool:
        """SELL3 storage (see the class notes); False, with nothing kept, when the
        field does not qualify."""
        dev = red.row_ptr.device
        n = self.n
        if n % 3:
            return False
        nf = n // 3
        # free DOFs in whole nodes: component c of the k-th free node is free DOF c nf + k
        free = red.free_idx.long()
        fk = free[:nf]
        if not (torch.equal(free[nf:2 * nf], fk + nper) and torch.equal(free[2 * nf:], fk + 2 * nper)):
            return False
        rp, col = red.row_ptr.long(), red.col.long()
        l3 = rp[1:] - rp[:-1]
        if bool((l3 % 3).any()) or not (torch.equal(l3[nf:2 * nf], l3[:nf]) and torch.equal(l3[2 * nf:], l3[:nf])):
            return False
        lk = l3[:nf] // 3
        nptr = torch.zeros(nf + 1, dtype=torch.int64, device=dev)
        nptr[1:] = torch.cumsum(lk, 0)
        nnz_n = int(nptr[-1].item())
        if nnz_n == 0:
            return False
        rowk = torch.repeat_interleave(torch.arange(nf, device=dev), lk)
        el = torch.arange(nnz_n, device=dev) - nptr[rowk]
        ncol = col[rp[rowk] + el]
        if int(ncol.max().item()) >= nf:
            return False
        # row c nf + k holds, for each column component t in turn, t nf + (the node columns of k)
        for c in range(3):
            base_c = rp[c * nf + rowk] + el
            for t in range(3):
                if not torch.equal(col[base_c + t * lk[rowk]], t * nf + ncol):
                    return False
        nptr32, ncol32 = nptr.to(torch.int32), ncol.to(torch.int32)
        perm_n = torch.empty(nf, dtype=torch.int32, device=dev)
        call("mife_sell_permutation", nf, ptr(nptr32), None, int(sigma), ptr(perm_n), _stream())
        pn = perm_n.long()
        iperm_n = torch.empty_like(perm_n)
        iperm_n[pn] = torch.arange(nf, dtype=torch.int32, device=dev)
        ns = (nf + 63) // 64
        lp = torch.zeros(ns * 64, dtype=torch.int64, device=dev)
        lp[:nf] = lk[pn]
        sp = torch.zeros(ns + 1, dtype=torch.int64, device=dev)
        sp[1:] = torch.cumsum(lp.view(ns, 64).amax(dim=1) * 64, 0)
        nblk = int(sp[-1].item())
        if 9 * nblk >= 2 ** 31:
            return False
        slice_ptr = sp.to(torch.int32)
        bcol = torch.empty(nblk, dtype=torch.int32, device=dev)
        bsrc = torch.empty(nblk, dtype=torch.int32, device=dev)
        call("mife_sell_build", nf, ptr(nptr32), ptr(ncol32), None, ptr(perm_n), ptr(iperm_n), ns, ptr(slice_ptr),
             ptr(bcol), ptr(bsrc), _stream())
        col16 = torch.empty(nblk, dtype=torch.int16, device=dev)
        cbase = torch.empty(ns, dtype=torch.int32, device=dev)
        ok = C.c_int(0)
        call("mife_sell_compress", ns, ptr(slice_ptr), ptr(bcol), ptr(cbase), ptr(col16), C.byref(ok), _stream())
        if not ok.value:
            return False
        # value t = 3 r + c of block b (node row k, node column at position e of
        # row k): solver entry (r nf + k, c nf + column), stored t-major per 64-block chunk
        q = bsrc.long()
        valid = q >= 0
        qv = q.clamp(min=0)
        kk = rowk[qv]
        ek = qv - nptr[kk]
        b = torch.arange(nblk, device=dev)
        lane = b % 64
        dst0 = (b - lane) * 9 + lane
        src_pos = red.src_pos.long()
        vsrc = torch.full((9 * nblk,), -1, dtype=torch.int32, device=dev)
        for t in range(9):
            r, c = divmod(t, 3)
            pos = rp[r * nf + kk] + c * lk[kk] + ek
            vsrc[dst0 + 64 * t] = torch.where(valid, src_pos[pos], -1).to(torch.int32)
        # vector position 3 k + c is solver row c nf + perm_n[k]
        perm = (torch.arange(3, device=dev).view(1, 3) * nf + pn.view(-1, 1)).reshape(-1).to(torch.int32)
        self.perm = perm
        self.iperm = torch.empty_like(perm)
        self.iperm[perm.long()] = torch.arange(n, dtype=torch.int32, device=dev)
        self.slice_ptr, self.col, self.col16, self.col_base, self.src = slice_ptr, bcol, col16, cbase, vsrc
        self.stored = 9 * nblk
        self.padding = self.stored / max(red.nnz, 1) - 1.0
        self.block = 3
        self.fmt = "sell3"
        self.slice_order = None
        span = C.c_int(0)
        if schedule:
            self.slice_order = torch.empty(_lib.lib().mife_sell_order_len(ns), dtype=torch.int32, device=dev)
            call("mife_sell_order", ns, ptr(slice_ptr), ptr(cbase), ptr(self.slice_order), C.byref(span), _stream())
        self.desc = MatrixDesc(n, _lib.MIFE_FMT_SELL3, ptr(slice_ptr), ptr(col16), ptr(perm), ns, self.stored,
                               ptr(cbase), ptr(self.slice_order), span.value)
        return True

    def gather(self, vals: torch.Tensor, prec=F64, out: torch.Tensor | None = None) -> torch.Tensor:
        """Solver-matrix values (this storage) from equations-CSR values (FP64 or FP32)."""
        if getattr(self.red, "coupled", False):
            # coupled rows and columns, or several sets: solver-CSR values first, then this storage
            v0 = vals[0] if isinstance(vals, (list, tuple)) else vals
            vals = self.red.gather(vals, F64 if v0.dtype == torch.float64 else F32)
            if self.fmt == "csr" and out is None and (vals.dtype == torch.float64) == (prec == F64):
                return vals
        if out is None:
            out = torch.empty(max(self.stored, 1), dtype=torch.float64 if prec == F64 else torch.float32,
                              device=vals.device)
        src_prec = F64 if vals.dtype == torch.float64 else F32
        call("mife_gather_values", self.stored, ptr(self.src), ptr(vals), src_prec, ptr(out), prec, _stream())
        return out

    def gather2(self, vals64: torch.Tensor, out64: torch.Tensor | None = None,
                out32: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        """FP64 solver-matrix values and the same values rounded to FP32, from FP64
        equations-CSR values in one pass (the mixed-precision solve's two copies)."""
        if getattr(self.red, "coupled", False):
            vals64 = self.red.gather(vals64, F64)
        n = max(self.stored, 1)
        if out64 is None:
            out64 = torch.empty(n, dtype=torch.float64, device=vals64.device)
        if out32 is None:
            out32 = torch.empty(n, dtype=torch.float32, device=vals64.device)
        call("mife_gather_values2", self.stored, ptr(self.src), ptr(vals64), ptr(out64), ptr(out32), _stream())
        return out64, out32

    def spmv(self, vals, x, y, prec=F64):
        """y = A x in this storage's row numbering (SELL: slice numbering)."""
        call("mife_spmv", C.byref(self.desc), ptr(vals), ptr(x), ptr(y), prec, _stream())
        return y


# Smallest solver matrix (stored entries) on which FP32 iterations pay for a scalar
# field: below it a solve is bound by launches and the Infinity Cache rather than by
# HBM bytes, so halving the value stream saves nothing.
MIXED_MIN_STORED = 16_000_000


def auto_precision(n_components: int, stored: int) -> str:
    """The precision "auto" picks for one solve: "mixed" (FP32 Jacobi CG with FP64
    residual replacement) for a scalar field whose solver matrix holds at least
    MIXED_MIN_STORED stored entries, else "fp64".  Measured on MI355X, full passes:
    48^3 triquadratic diffusion (54 M entries) mixed 36 ms against 51 ms in FP64;
    16^3 trilinear Laplace (0.1 M) 2.0 ms against 0.4 ms, where launches rather
    than bytes bound an iteration.  A three-component field (96x32x32 trilinear
    elasticity, 3x3 node blocks) runs about 50% more iterations on the rounded
    FP32 operator than FP64 CG (1,570-1,870 against 1,212, whatever the inner
    vectors' precision), which eats the halved value stream: mixed 62-69 ms
    against FP64 64-69 ms, so vector fields stay in FP64."""
    return "mixed" if n_components == 1 and stored >= MIXED_MIN_STORED else "fp64"


@dataclass
class SolveInfo:
    iterations: int
    relres: float
    outer: int = 0


class KrylovWorkspace:
    """Device workspace of the Krylov solvers on n rows; GMRES restarts up to
    ``restart`` fit (``ensure`` grows it for a longer restart)."""

    def __init__(self, n: int, device, restart: int = 30):
        self.n = n
        self.device = device
        self.restart = 0
        self.buf = None
        self.ensure(restart)

    def ensure(self, restart: int):
        restart = max(int(restart), 1)
        if self.buf is None or restart > self.restart:
            self.restart = restart
            self.nbytes = int(_lib.lib().mife_solver_work_bytes(self.n, restart))
            self.buf = None
            self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=self.device)
        return self


def _call_solver(name, *args, it, rr, outer=None):
    try:
        call(name, *args)
    except _lib.NotConvergedError as e:
        e.info = SolveInfo(it.value, rr.value, outer.value if outer is not None else 0)
        raise
    return SolveInfo(it.value, rr.value, outer.value if outer is not None else 0)


def cg(A: SolverMatrix, vals, b, x, rtol, maxit, ws: KrylovWorkspace, prec=F64) -> SolveInfo:
    """Jacobi CG on A (b, x in the original free-DOF numbering).  On reaching maxit
    raises NotConvergedError with ``.info`` (iterations, relative residual)."""
    it, rr = C.c_int(0), C.c_double(0.0)
    return _call_solver("mife_cg", C.byref(A.desc), ptr(vals), ptr(b), ptr(x), prec, float(rtol), int(maxit),
                        ptr(ws.buf), ws.nbytes, C.byref(it), C.byref(rr), _stream(), it=it, rr=rr)


def refine(A: SolverMatrix, vals64, vals32, b, x, rtol, maxit, ws: KrylovWorkspace, inner_rtol=0.1) -> SolveInfo:
    """FP32 Jacobi CG with FP64 residual replacement every ``inner_rtol`` reduction."""
    it, outer, rr = C.c_int(0), C.c_int(0), C.c_double(0.0)
    return _call_solver("mife_refine", C.byref(A.desc), ptr(vals64), ptr(vals32), ptr(b), ptr(x), float(rtol),
                        float(inner_rtol), int(maxit), ptr(ws.buf), ws.nbytes, C.byref(it), C.byref(outer),
                        C.byref(rr), _stream(), it=it, rr=rr, outer=outer)


def gmres(A: SolverMatrix, vals, b, x, rtol, maxit, ws: KrylovWorkspace, restart=30, pc_inv=None) -> SolveInfo:
    """Restarted GMRES; pc_inv (FP64, solver-row order) overrides the Jacobi entry
    of rows where it is nonzero.  The workspace grows to fit ``restart``."""
    ws.ensure(restart)
    it, rr = C.c_int(0), C.c_double(0.0)
    return _call_solver("mife_gmres", C.byref(A.desc), ptr(vals), ptr(b), ptr(x), ptr(pc_inv), float(rtol),
                        int(restart), int(maxit), ptr(ws.buf), ws.nbytes, C.byref(it), C.byref(rr), _stream(),
                        it=it, rr=rr)


def minres(A: SolverMatrix, vals, b, x, rtol, maxit, ws: KrylovWorkspace, pc_inv=None) -> SolveInfo:
    """Preconditioned MINRES for a symmetric (indefinite) solver matrix, FP64 vectors;
    vals FP64 or FP32 (read into FP64 arithmetic).  rtol applies to the preconditioned
    residual norm, SolveInfo.relres is the true relative residual of that operator."""
    it, rr = C.c_int(0), C.c_double(0.0)
    prec = F32 if vals.dtype == torch.float32 else F64
    return _call_solver("mife_minres", C.byref(A.desc), ptr(vals), prec, ptr(b), ptr(x), ptr(pc_inv), float(rtol),
                        int(maxit), ptr(ws.buf), ws.nbytes, C.byref(it), C.byref(rr), _stream(), it=it, rr=rr)


def backsubstitute(csr: CSRStructure, vals64, u, src=None, rhs_fixed=None, iron_compat=False, out=None):
    if out is None:
        out = torch.empty(csr.n, dtype=torch.float64, device=vals64.device)
    call("mife_backsubstitute", csr.n, ptr(csr.row_ptr), ptr(csr.col), ptr(vals64), ptr(u), ptr(src),
         ptr(rhs_fixed), 1 if iron_compat else 0, ptr(out), _stream())
    return out


def grid_element_dofs(grid, P: int, n_comp: int) -> np.ndarray:
    """[E][n_comp (P+1)^3] DOFs of a generated regular grid of order P: elements and
    nodes numbered x fastest, local nodes xi1 fastest, components in separated blocks
    of one DOF per node (generated_mesh_routines.F90:1568-1612)."""
    nx, ny, nz = (int(g) for g in grid)
    NB = P + 1
    Nx, Ny, Nz = P * nx + 1, P * ny + 1, P * nz + 1
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    az, ay, ax = np.meshgrid(np.arange(NB), np.arange(NB), np.arange(NB), indexing="ij")
    node = ((P * ex.reshape(-1, 1) + ax.reshape(1, -1)) + Nx * ((P * ey.reshape(-1, 1) + ay.reshape(1, -1))
                                                                + Ny * (P * ez.reshape(-1, 1) + az.reshape(1, -1))))
    return np.concatenate([c * Nx * Ny * Nz + node for c in range(n_comp)], axis=1).astype(np.int32)


class _GridStage:
    """Structured assembly on a generated regular grid (mife_assemble_grid): element
    matrices staged in one launch, then every CSR row written whole.  Holds the stage
    buffer (sized for FP64, shared by both precisions)."""

    def __init__(self, grid, basis: LagrangeBasis, n_comp: int, csr: CSRStructure):
        self.grid = np.ascontiguousarray(grid, dtype=np.int32)
        self.n_comp = n_comp
        self.csr = csr
        E = int(np.prod(self.grid))
        # the kernels derive every row and element from grid indices: the element DOFs
        # must be the generated mesh's (nodes x fastest, components separated)
        want = grid_element_dofs(self.grid, 1 if basis.num_nodes == 8 else 2, n_comp)
        if csr.n != int(want.max()) + 1 or not np.array_equal(csr.elem_dofs.cpu().numpy(), want):
            raise ValueError(f"the element DOFs are not those of a generated {tuple(self.grid)} grid")
        self.nbytes = int(_lib.lib().mife_assemble_grid_stage_bytes(E, basis.num_nodes, n_comp, F64))
        self.buf = None

    def run(self, desc, coords, enodes, coef, source, prec, vals, src_vec):
        if self.buf is None:
            self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=self.csr.device)
        call("mife_assemble_grid", C.byref(desc), self.grid.ctypes.data, self.n_comp, ptr(coords), ptr(enodes),
             ptr(coef), ptr(source), ptr(self.csr.row_ptr), self.csr.n, prec, ptr(vals), ptr(src_vec),
             ptr(self.buf), self.nbytes, _stream())


def _grid_ok(grid, basis: LagrangeBasis) -> bool:
    return grid is not None and basis.num_xi == 3 and (basis.num_nodes, basis.num_gauss) in ((8, 8), (27, 27))


class ScalarAssembler:
    """Laplace / generalised Laplace / constant-source Poisson on one basis.

    ``grid`` (elements per direction of a generated regular mesh, numbered as iron's
    generator numbers them) selects the structured assembly, which writes the CSR
    values and the source vector whole (bit for bit the coloured scatter's sums);
    otherwise the element kernels add into ``vals`` through the position map."""

    def __init__(self, basis: LagrangeBasis, coords: np.ndarray, elem_nodes: np.ndarray, csr: CSRStructure,
                 grid=None):
        dev = csr.device
        self.csr = csr
        self.db = DeviceBasis(basis, dev)
        self.coords = torch.as_tensor(np.ascontiguousarray(coords, dtype=np.float64), device=dev)
        self.elem_nodes = torch.as_tensor(np.ascontiguousarray(elem_nodes, dtype=np.int32), device=dev)
        self.n_elem = elem_nodes.shape[0]
        self.grid = _GridStage(grid, basis, 1, csr) if _grid_ok(grid, basis) else None

    def assemble(self, vals, kappa=None, source=None, src_vec=None, prec=F64, atomic=False):
        cs = self.csr
        if self.grid is not None and not atomic:
            self.grid.run(self.db.desc, self.coords, self.elem_nodes, kappa, source, prec, vals, src_vec)
            return vals
        call("mife_assemble_scalar", C.byref(self.db.desc), self.n_elem, ptr(self.coords), ptr(self.elem_nodes),
             ptr(cs.elem_dofs), ptr(kappa), ptr(source), ptr(cs.pos), None if atomic else ptr(cs.order),
             None if atomic else cs.colour_ptr.ctypes.data, cs.n_colours, prec, ptr(vals), ptr(src_vec), _stream())
        return vals


class ElasticAssembler:
    """3-D linear elasticity on one basis (all three components on it); ``grid`` as
    for ScalarAssembler."""

    def __init__(self, basis: LagrangeBasis, coords: np.ndarray, elem_nodes: np.ndarray, csr: CSRStructure,
                 grid=None):
        dev = csr.device
        self.csr = csr
        self.db = DeviceBasis(basis, dev)
        self.coords = torch.as_tensor(np.ascontiguousarray(coords, dtype=np.float64), device=dev)
        self.elem_nodes = torch.as_tensor(np.ascontiguousarray(elem_nodes, dtype=np.int32), device=dev)
        self.n_elem = elem_nodes.shape[0]
        ok = _grid_ok(grid, basis) and basis.num_nodes == 8
        self.grid = _GridStage(grid, basis, 3, csr) if ok else None

    def assemble(self, vals, material, prec=F64, atomic=False):
        cs = self.csr
        if self.grid is not None and not atomic:
            self.grid.run(self.db.desc, self.coords, self.elem_nodes, material, None, prec, vals, None)
            return vals
        call("mife_assemble_elastic", C.byref(self.db.desc), self.n_elem, ptr(self.coords), ptr(self.elem_nodes),
             ptr(cs.elem_dofs), ptr(material), ptr(cs.pos), None if atomic else ptr(cs.order),
             None if atomic else cs.colour_ptr.ctypes.data, cs.n_colours, prec, ptr(vals), _stream())
        return vals
