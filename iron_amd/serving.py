"""Serving many small finite-element solves on one GPU.

Requests (a diffusion or an elasticity box of its own size and coefficients)
arrive over time.  Each admitted request is assembled and reduced on the GPU
(the same kernels as a single solve), and its reduced system is stored into
shared HBM pools in SELL-64 form (slice pointers, 16-bit or int32 columns,
values; vectors in slice numbering), next to the CG vectors.
All admitted requests then iterate together in grouped Jacobi CG launches: one
launch serves every unconverged system, each with its own scalars and stopping
test.  A request iterates either in FP64 or in FP32 (FP32 copy of the values,
FP32 vectors) with FP64 reliable updates: its solution is accumulated in FP64
and its residual is replaced by the true FP64 residual b - A x after every
decade of FP32 reduction, so it stops on its FP64 residual (1e-8 relative).
Both kinds share the same launches.  The default policy ("auto") applies the
single-solve rule per request, which keeps these boxes (at most a few million
stored entries) in FP64: there a request is bound by launches and the Infinity
Cache, not by HBM bytes, so FP32 values save little, and on the lognormal
elasticity boxes the FP32 operator's rounding (a 1e-4..1e-3 relative change of
A^-1) makes each replacement restart part of the convergence -- about 2.8x the
FP64 iterations over a 1024-request stream, a net loss ("mixed" runs it; "split"
runs diffusion in FP32 and elasticity in FP64).  Requests are stored on
device-built box structures (mife_box_build) and assembled by the structured
grid path.  Between chunks of iterations the server
retires converged requests (their solution goes back to the host and their
pool ranges are freed) and admits waiting ones, first come first served, as
long as their rows and entries fit the pool budget (HBM-budget admission).  A
chunk of iterations of an unchanged group is captured once in a HIP graph and
replayed.

No reference counterpart: iron solves one problem per Problem_Solve.  A
request's iterates depend only on its own data, never on which other requests
share the launches (tests compare every request against the FP64 oracle, and a
request solved alone against the same request solved in a mix, bit for bit).
"""
from __future__ import annotations

import ctypes as C
import time
from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, fem
from ._lib import Pools, Segment, call, ptr
from .basis import basis



@dataclass
class Request:
    rid: int
    kind: str                 # "diffusion" or "elasticity"
    n_elem: tuple
    seed: int
    arrival: float = 0.0      # seconds after the start of the run
    rtol: float = 1e-8
    # results
    u: np.ndarray | None = None
    iterations: int = 0
    relres: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


# ------------------------------------------------------------------ problems
def lognormal_field(n: int, seed: int, device) -> torch.Tensor:
    """exp(N(0,1)) per element, drawn on the device from the request's seed (the
    device's Philox generator: no host draw or upload per request)."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return torch.randn(n, dtype=torch.float64, generator=g, device=device).exp_()


_Q1 = {}


def _q1(device) -> fem.DeviceBasis:
    """Trilinear basis tables on `device` (shared by every box)."""
    key = str(device)
    if key not in _Q1:
        _Q1[key] = fem.DeviceBasis(basis(3, 1, 2), device)
    return _Q1[key]


class _BoxSolverMatrix:
    """The reduced matrix's storage of a box (solver order, uniform slice width):
    SELL-64, or SELL3 node blocks for an elasticity box (free DOFs node by node).
    What the grouped server copies into its pools, and a solver matrix of its own
    for the single-solve path (identity permutation)."""

    def __init__(self, cnt: _lib.BoxCounts, device):
        self.n = int(cnt.n_free)
        self.stored = int(cnt.stored)
        self.block = 3 if cnt.block else 1
        ns = int(cnt.n_slices)
        self.slice_ptr = torch.empty(ns + 1, dtype=torch.int32, device=device)
        self.src = torch.empty(self.stored, dtype=torch.int32, device=device)
        # SELL3: one 16-bit node column per block of nine values
        self.n_col = self.stored // self.block
        if cnt.col16:
            self.col16 = torch.empty(self.n_col, dtype=torch.int16, device=device)
            self.col_base = torch.empty(ns, dtype=torch.int32, device=device)
            self.col = None
        else:
            self.col = torch.empty(self.stored, dtype=torch.int32, device=device)
            self.col16 = self.col_base = None
        self.perm = torch.arange(self.n, dtype=torch.int32, device=device)
        colp = ptr(self.col16) if self.col16 is not None else ptr(self.col)
        self.format = _lib.MIFE_FMT_SELL3 if self.block == 3 else _lib.MIFE_FMT_SELL
        self.desc = _lib.MatrixDesc(self.n, self.format, ptr(self.slice_ptr), colp, ptr(self.perm), ns,
                                    self.stored, ptr(self.col_base), None, 0)

    def gather(self, vals, prec=fem.F64, out=None):
        if out is None:
            out = torch.empty(self.stored, dtype=torch.float64 if prec == fem.F64 else torch.float32,
                              device=vals.device)
        call("mife_gather_values", self.stored, ptr(self.src), ptr(vals), fem.F64, ptr(out), prec, fem._stream())
        return out


class _BoxReduction:
    """Dirichlet reduction of a box: free DOFs ascending (the builder's new_idx /
    free_idx), the right-hand side b = f - K_fd u_d and the field update."""

    def __init__(self, st: "_Structure", cnt: _lib.BoxCounts, device):
        self.st = st
        self.n_free = int(cnt.n_free)
        self.nnz = int(cnt.red_nnz)
        self.free_idx = torch.empty(self.n_free, dtype=torch.int32, device=device)
        self.new_idx = torch.empty(int(cnt.n_dofs), dtype=torch.int32, device=device)

    def rhs(self, vals, u, rhs=None, out=None):
        st = self.st
        if out is None:
            out = torch.empty(self.n_free, dtype=torch.float64, device=vals.device)
        call("mife_solver_rhs", self.n_free, ptr(self.free_idx), ptr(st.row_ptr), ptr(st.col), ptr(vals),
             ptr(self.new_idx), ptr(u), None, ptr(rhs), None, ptr(out), fem._stream())
        return out

    def field_update(self, x, u):
        call("mife_field_update", self.n_free, ptr(self.free_idx), ptr(x), ptr(u), fem._stream())


class _Structure:
    """Device structures of one request shape (cached by (kind, n_elem)), built by
    mife_box_build from closed forms of the box: trilinear elements, the
    equations CSR, the Dirichlet reduction (diffusion: u = 0 on x = 0 and u = 1 on
    x = 1; elasticity: clamped at x = 0) and the reduced matrix's storage (SELL-64;
    elasticity boxes SELL3 node blocks while ``block`` is set).  The sizes come from
    the box alone, so nothing waits for the device."""

    block = True

    def __init__(self, kind: str, n_elem, device):
        self.kind = kind
        self.n_elem = tuple(int(n) for n in n_elem)
        self.device = device
        ncomp = 1 if kind == "diffusion" else 3
        self.ncomp = ncomp
        extent = (1.0, 1.0, 1.0) if kind == "diffusion" else (3.0, 1.0, 1.0)
        desc = _lib.BoxDesc(self.n_elem, 1, ncomp, 1, 1 if kind == "diffusion" else 0,
                            1 if ncomp == 3 and self.block else 0, extent, 0.0, 1.0 if kind == "diffusion" else 0.0)
        cnt = _lib.BoxCounts()
        call("mife_box_sizes", C.byref(desc), C.byref(cnt))
        self.num_elements = int(cnt.n_elem)
        self.num_nodes = int(cnt.n_nodes)
        self.n = int(cnt.n_dofs)
        self.nnz = int(cnt.nnz)
        dev = device
        self.coords = torch.empty((self.num_nodes, 3), dtype=torch.float64, device=dev)
        self.elem_nodes = torch.empty((self.num_elements, 8), dtype=torch.int32, device=dev)
        self.row_ptr = torch.empty(self.n + 1, dtype=torch.int32, device=dev)
        self.col = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        self.u0 = torch.empty(self.n, dtype=torch.float64, device=dev)
        self.red = _BoxReduction(self, cnt, dev)
        self.S = _BoxSolverMatrix(cnt, dev)
        S = self.S
        arrays = _lib.BoxArrays(ptr(self.coords), ptr(self.elem_nodes), ptr(self.row_ptr), ptr(self.col),
                                ptr(self.red.free_idx), ptr(self.red.new_idx), ptr(self.u0), ptr(S.slice_ptr),
                                ptr(S.col), ptr(S.col16), ptr(S.col_base), ptr(S.src))
        call("mife_box_build", C.byref(desc), C.byref(arrays), fem._stream())
        self.rhs = None
        if kind == "elasticity":
            # consistent nodal forces of t_z = -1 on the x = 3 face: int phi dA on the
            # bilinear face grid is h_y h_z (1/2 at face edges, 1 inside) per direction
            ny, nz = self.n_elem[1], self.n_elem[2]
            Nx, Ny = self.n_elem[0] + 1, ny + 1
            iy = torch.arange(ny + 1, device=dev)
            iz = torch.arange(nz + 1, device=dev)
            cy = torch.where((iy == 0) | (iy == ny), 0.5, 1.0).to(torch.float64)
            cz = torch.where((iz == 0) | (iz == nz), 0.5, 1.0).to(torch.float64)
            w = (-1.0 * (1.0 / ny) * (1.0 / nz)) * (cz[:, None] * cy[None, :])
            tip = (Nx - 1) + Nx * (iy[None, :] + Ny * iz[:, None])
            self.rhs = torch.zeros(self.n, dtype=torch.float64, device=dev)
            self.rhs[2 * self.num_nodes + tip.reshape(-1)] = w.reshape(-1)
        self.stage_bytes = int(_lib.lib().mife_assemble_grid_stage_bytes(self.num_elements, 8, ncomp, fem.F64))
        self._grid = np.array(self.n_elem, dtype=np.int32)

    def zeros(self):
        return torch.zeros(self.nnz, dtype=torch.float64, device=self.device)

    def coefficients(self, seed: int, device):
        """Lognormal(0,1) element coefficients from the request's seed: conductivity, or
        Young's modulus 10 exp(N(0,1)) with Poisson ratio 0.3 (material [E][6])."""
        z = lognormal_field(self.num_elements, seed, device)
        if self.kind == "diffusion":
            return z
        Ey = 10.0 * z
        nu = torch.full_like(Ey, 0.3)
        return torch.stack([Ey, Ey, Ey, nu, nu, nu], dim=1).contiguous()

    def assemble(self, coef, vals, stage):
        """FP64 equations values (overwritten) by the structured assembly; `stage`
        holds at least stage_bytes bytes."""
        q1 = _q1(self.device)
        call("mife_assemble_grid", C.byref(q1.desc), self._grid.ctypes.data, self.ncomp, ptr(self.coords),
             ptr(self.elem_nodes), ptr(coef), None, ptr(self.row_ptr), self.n, fem.F64, ptr(vals), None,
             ptr(stage), stage.numel(), fem._stream())
        return vals


# ------------------------------------------------------------------ pools
class _FirstFit:
    def __init__(self, cap: int):
        self.cap = int(cap)
        self.free = [(0, self.cap)]  # sorted (start, size)

    def alloc(self, size: int):
        size = max(int(size), 1)
        for i, (s, n) in enumerate(self.free):
            if n >= size:
                if n == size:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + size, n - size)
                return s
        return None

    def release(self, start: int, size: int):
        size = max(int(size), 1)
        self.free.append((start, size))
        self.free.sort()
        merged = []
        for s, n in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + n)
            else:
                merged.append((s, n))
        self.free = merged


class Server:
    """Grouped-CG serving of Request objects within an HBM budget (bytes).

    precision: "auto" (per request, the single-solve rule fem.auto_precision for scalar
    fields: FP32 iterations with FP64 reliable updates when the solver matrix
    holds at least fem.MIXED_MIN_STORED entries, else FP64 -- on boxes of up to 64^3
    trilinear elements every request iterates in FP64), "split" (FP32 + reliable
    updates for diffusion requests, FP64 for elasticity), "mixed" (every request
    FP32 + reliable updates) or "fp64".  Reliable updates follow every
    ``inner_rtol`` reduction of a request's FP32 residual, tested every
    ``check_every`` iterations."""

    MIXED_KINDS = {"auto": None, "split": ("diffusion",), "mixed": ("diffusion", "elasticity"), "fp64": ()}

    def __init__(self, hbm_budget: float = 32e9, max_segments: int = 256, nz_per_row: float = 60.0,
                 chunk: int = 16, use_graph: bool = True, device=None, cache_shapes: int = 8,
                 admit_budget: float = 0.004, precision: str = "auto", inner_rtol: float = 0.1,
                 check_every: int = 4, prefetch: int = 4):
        self.device = fem._dev(device)
        self.max_segments = int(max_segments)
        if precision not in self.MIXED_KINDS:
            raise ValueError(f"precision must be one of {sorted(self.MIXED_KINDS)}")
        self.mixed_kinds = self.MIXED_KINDS[precision]
        self.inner_rtol = float(inner_rtol)
        self.check_every = int(check_every)
        # pool bytes per row: b, x (FP64), D^-1, r, z, p, q (FP64), D^-1, r, z, p, q, x32 (FP32)
        # and the slice pointer / base column share; per stored entry: FP64 value, FP32
        # copy, int32 column, 16-bit column (all pools exist)
        self.bytes_per_row = 2 * 8 + 5 * 8 + 6 * 4 + 4
        self.bytes_per_nz = 8 + 4 + 4 + 2
        rows = int(hbm_budget / (self.bytes_per_row + nz_per_row * self.bytes_per_nz))
        self.rows_cap = max(rows, 1024)
        self.nz_cap = int(self.rows_cap * nz_per_row)
        self.rp_cap = self.rows_cap + self.max_segments
        dev = self.device
        self.b = torch.zeros(self.rows_cap, dtype=torch.float64, device=dev)
        self.x = torch.zeros(self.rows_cap, dtype=torch.float64, device=dev)
        self.vec5 = torch.zeros(5 * self.rows_cap, dtype=torch.float64, device=dev)
        self.vec6 = torch.zeros(6 * self.rows_cap, dtype=torch.float32, device=dev)
        self.rp = torch.zeros(self.rp_cap, dtype=torch.int32, device=dev)
        self.col = torch.zeros(max(self.nz_cap, 1), dtype=torch.int32, device=dev)
        self.col16 = torch.zeros(max(self.nz_cap, 1), dtype=torch.int16, device=dev)
        self.cb_cap = self.rows_cap // 64 + 2 * self.max_segments
        self.cbase = torch.zeros(self.cb_cap, dtype=torch.int32, device=dev)
        self.val = torch.zeros(max(self.nz_cap, 1), dtype=torch.float64, device=dev)
        self.val32 = torch.zeros(max(self.nz_cap, 1), dtype=torch.float32, device=dev)
        self.pools = Pools(ptr(self.rp), ptr(self.col), ptr(self.col16), ptr(self.cbase), ptr(self.val),
                           ptr(self.val32))
        self.max_blocks = _lib.lib().mife_grouped_blocks(self.rows_cap) + 2 * self.max_segments
        L = _lib.lib()
        self.state = torch.zeros(int(L.mife_grouped_state_bytes(self.max_segments)), dtype=torch.uint8, device=dev)
        wb = int(L.mife_grouped_work_bytes(self.max_segments, self.max_blocks))
        self.table = torch.zeros(wb, dtype=torch.uint8, device=dev)
        self.init_table = torch.zeros(wb, dtype=torch.uint8, device=dev)
        self.resid_table = torch.zeros(wb, dtype=torch.uint8, device=dev)
        self.side = torch.cuda.Stream()
        self.staging = np.zeros(wb, dtype=np.uint8)
        self.alloc_rows = _FirstFit(self.rows_cap)
        self.alloc_rp = _FirstFit(self.rp_cap)
        self.alloc_nz = _FirstFit(self.nz_cap)
        self.alloc_cb = _FirstFit(self.cb_cap)
        self.free_slots = list(range(self.max_segments))[::-1]
        self.chunk = int(chunk)
        self.admit_budget = float(admit_budget)
        self.use_graph = bool(use_graph)
        self.cache = OrderedDict()
        self.cache_shapes = int(cache_shapes)
        self.graph = None
        self.stats = dict(admitted=0, retired=0, chunks=0, graph_captures=0, graph_replays=0, table_builds=0,
                          iterations=0, spmv_bytes=0.0, assembly_batches=0)
        # structures of the next `prefetch` queued requests are built by one worker
        # thread on its own stream while the main thread waits for the GPU
        self.prefetch = int(prefetch)
        self._pool = ThreadPoolExecutor(max_workers=1) if self.prefetch > 0 else None
        self._pf_stream = torch.cuda.Stream() if self.prefetch > 0 else None
        self._pending = {}
        self.stage = None   # element-matrix stage of the structured assembly (largest box so far)
        # host-side wall time per phase (seconds): structure builds, admission
        # (assembly + reduction + pool copies), iteration chunks, status reads, retirement
        self.timing = dict(structure=0.0, admit=0.0, chunk=0.0, status=0.0, retire=0.0)

    # -- structures
    def _build_structure(self, key):
        with torch.cuda.stream(self._pf_stream):
            st = _Structure(key[0], key[1], self.device)
        self._pf_stream.synchronize()   # complete before another stream reads it
        return st

    def _prefetch(self, queue, qi, now):
        """Queues structure builds for the next waiting requests that have arrived."""
        if self._pool is None:
            return
        for r in queue[qi:qi + self.prefetch]:
            if r.arrival > now:
                break
            key = (r.kind, tuple(r.n_elem))
            if key not in self.cache and key not in self._pending:
                self._pending[key] = self._pool.submit(self._build_structure, key)

    def _structure(self, req: Request) -> _Structure:
        key = (req.kind, tuple(req.n_elem))
        st = self.cache.get(key)
        if st is None:
            t = time.perf_counter()
            fut = self._pending.pop(key, None)
            st = fut.result() if fut is not None else _Structure(req.kind, req.n_elem, self.device)
            self.timing["structure"] += time.perf_counter() - t
            self.cache[key] = st
            while len(self.cache) > self.cache_shapes:
                self.cache.popitem(last=False)
        else:
            self.cache.move_to_end(key)
        return st

    def _mixed(self, st: "_Structure") -> bool:
        if self.mixed_kinds is None:
            # scalar fields only: the grouped solver's FP32 elasticity iterations are
            # not measured to pay
            return st.ncomp == 1 and fem.auto_precision(st.ncomp, st.S.stored) == "mixed"
        return st.kind in self.mixed_kinds

    def fits(self, req: Request) -> bool:
        st = self._structure(req)
        S = st.S
        ns = S.desc.n_slices
        return (self._largest(self.alloc_rows) >= S.n and self._largest(self.alloc_nz) >= S.stored
                and self._largest(self.alloc_rp) >= ns + 1 and self._largest(self.alloc_cb) >= ns
                and bool(self.free_slots))

    @staticmethod
    def _largest(a: _FirstFit) -> int:
        return max((n for _, n in a.free), default=0)

    # -- admission: assemble, reduce, store into the pools
    def admit(self, req: Request, now: float):
        """Admits one request: pool ranges, assembly, reduction into the pools."""
        e = self._prepare(req, now)
        if e is not None:
            self._assemble_batch([e])
            self._finish(e)
        return e

    def _prepare(self, req: Request, now: float):
        """Pool ranges, a slot and zeroed equations values for one request (None if
        it does not fit)."""
        st = self._structure(req)
        S = st.S
        nf, stored, ns = S.n, S.stored, S.desc.n_slices
        want = ((self.alloc_rows, nf), (self.alloc_rp, ns + 1), (self.alloc_nz, stored), (self.alloc_cb, ns))
        offs = [a.alloc(n) for a, n in want]
        if any(o is None for o in offs):
            for (a, n), o in zip(want, offs):
                if o is not None:
                    a.release(o, n)
            return None
        ro, rpo, nzo, cbo = offs
        slot = self.free_slots.pop()
        coef = st.coefficients(req.seed, self.device)
        req.t_admit = now
        self.stats["admitted"] += 1
        return dict(req=req, st=st, nf=nf, nnz=st.red.nnz, stored=stored, ns=ns, ro=ro, rpo=rpo, nzo=nzo, cbo=cbo,
                    slot=slot, coef=coef, vals=torch.empty(st.nnz, dtype=torch.float64, device=self.device),
                    mixed=self._mixed(st))

    def _assemble_batch(self, entries):
        """FP64 equations values of every entry: the structured assembly of each box
        (element matrices staged, then every CSR row written whole), one stage
        buffer shared by the admissions, which run in order on one stream."""
        need = max(e["st"].stage_bytes for e in entries)
        if self.stage is None or self.stage.numel() < need:
            self.stage = torch.empty(need, dtype=torch.uint8, device=self.device)
        for e in entries:
            e["st"].assemble(e["coef"], e["vals"], self.stage)
        self.stats["assembly_batches"] += 1

    def _finish(self, e):
        """Solver values, slice structure and right-hand side of an assembled entry
        into the pools; its segment."""
        st, vals = e["st"], e.pop("vals")
        e.pop("coef")
        red, S = st.red, st.S
        nf, stored, ns = e["nf"], e["stored"], e["ns"]
        ro, rpo, nzo, cbo = e["ro"], e["rpo"], e["nzo"], e["cbo"]
        c16 = S.col16 is not None
        s = fem._stream()
        # solver values straight from the equations CSR into the value pool
        call("mife_gather_values", stored, ptr(S.src), ptr(vals), fem.F64, ptr(self.val) + 8 * nzo, fem.F64, s)
        if e["mixed"]:   # the FP32 operator is the FP64 one rounded
            call("mife_gather_values", stored, ptr(S.src), ptr(vals), fem.F64, ptr(self.val32) + 4 * nzo, fem.F32, s)
        self.rp[rpo:rpo + ns + 1].copy_(S.slice_ptr[: ns + 1])
        if c16:
            self.col16[nzo:nzo + S.n_col].copy_(S.col16[:S.n_col])
            self.cbase[cbo:cbo + ns].copy_(S.col_base[:ns])
        else:
            self.col[nzo:nzo + stored].copy_(S.col[:stored])
        call("mife_solver_rhs", nf, ptr(red.free_idx), ptr(st.row_ptr), ptr(st.col), ptr(vals),
             ptr(red.new_idx), ptr(st.u0), None, ptr(st.rhs), None, ptr(self.b) + 8 * ro, s)
        self.x[ro:ro + nf].zero_()
        e["seg"] = Segment(nf, 200000, e["slot"], S.format, ro, rpo, nzo, e["req"].rtol, ns,
                           fem.F32 if e["mixed"] else fem.F64, cbo if c16 else -1)
        e["cbo_used"] = c16

    def _build_table(self, entries, table):
        segs = (Segment * max(len(entries), 1))(*[e["seg"] for e in entries])
        nb = C.c_int(0)
        call("mife_grouped_setup", segs, len(entries), self.max_segments, self.max_blocks, C.byref(nb),
             self.staging.ctypes.data, ptr(table), fem._stream())
        self.stats["table_builds"] += 1
        return nb.value

    def _iterate(self, n_seg, nb, iters):
        call("mife_grouped_cg_iterate", n_seg, nb, self.max_segments, self.max_blocks, C.byref(self.pools),
             ptr(self.b), ptr(self.x), ptr(self.vec5), ptr(self.vec6), self.rows_cap, int(iters), self.check_every,
             self.inner_rtol, ptr(self.table), ptr(self.state), fem._stream())

    def _init(self, n_seg, nb, table):
        call("mife_grouped_cg_init", n_seg, nb, self.max_segments, self.max_blocks, C.byref(self.pools),
             ptr(self.b), ptr(self.x), ptr(self.vec5), ptr(self.vec6), self.rows_cap, ptr(table), ptr(self.state),
             fem._stream())

    def _status(self):
        it = np.zeros(self.max_segments, dtype=np.int32)
        dn = np.zeros(self.max_segments, dtype=np.int32)
        rr = np.zeros(self.max_segments, dtype=np.float64)
        call("mife_grouped_cg_status", self.max_segments, ptr(self.state), it.ctypes.data, dn.ctypes.data,
             rr.ctypes.data, fem._stream())
        return it, dn, rr

    def _retire(self, e, now):
        req, seg, st = e["req"], e["seg"], e["st"]
        nf = e["nf"]
        u = st.u0.clone()
        st.red.field_update(self.x[seg.row_off:seg.row_off + nf], u)
        req.u = u.cpu().numpy()
        req.t_done = time.perf_counter() - now
        self.alloc_rows.release(seg.row_off, nf)
        self.alloc_rp.release(seg.rp_off, e["ns"] + 1)
        self.alloc_nz.release(seg.nz_off, e["stored"])
        self.alloc_cb.release(e["cbo"], e["ns"])
        self.free_slots.append(seg.slot)
        self.stats["retired"] += 1

    # -- the serving loop
    def run(self, requests, max_batch: int | None = None, overlap: bool = True):
        """Serve ``requests`` (arrival times relative to the call); returns them with
        results, iterations, relative residuals and completion times filled in.

        With ``overlap`` the assembly, reduction and pool set-up of arriving requests
        run on a second stream while the group's iteration chunk runs on the main
        stream; they join the group at the next chunk boundary."""
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active, pending = [], []
        t0 = time.perf_counter()
        qi = 0
        nb = 0
        changed = True
        main = torch.cuda.current_stream()
        side = self.side if overlap else main
        while qi < len(queue) or active or pending:
            self._prefetch(queue, qi, time.perf_counter() - t0)
            if pending:
                main.wait_stream(side)
                active += pending
                pending = []
                changed = True
            launched = False
            if active:
                if changed:
                    nb = self._build_table(active, self.table)
                    self.graph = None
                    changed = False
                tc = time.perf_counter()
                self._run_chunk(len(active), nb)
                self.stats["chunks"] += 1
                self.timing["chunk"] += time.perf_counter() - tc
                launched = True
            # admission of arrived requests (FIFO, stop at the first that does not fit).
            # The host work of an admission (structure, assembly, pool copies) is
            # paced against the GPU: while a chunk runs, admissions stop once they
            # have taken admit_budget seconds, and with nothing running the first
            # admitted request starts iterating at once
            ta = time.perf_counter()
            with torch.cuda.stream(side):
                new = []
                now = time.perf_counter() - t0
                while qi < len(queue) and queue[qi].arrival <= now and (
                        max_batch is None or len(active) + len(new) < max_batch):
                    if new and (not (launched or pending) or time.perf_counter() - ta > self.admit_budget):
                        break
                    e = self._prepare(queue[qi], now) if self.fits(queue[qi]) else None
                    if e is None:
                        break
                    new.append(e)
                    qi += 1
                if new:
                    # the arrivals of this round: one batched assembly, then each into the pools
                    self._assemble_batch(new)
                    for e in new:
                        self._finish(e)
                    nbi = self._build_table(new, self.init_table)
                    self._init(len(new), nbi, self.init_table)
                    pending += new
            self.timing["admit"] += time.perf_counter() - ta
            if not launched:
                if not pending and qi < len(queue):
                    time.sleep(max(0.0, min(queue[qi].arrival - (time.perf_counter() - t0), 0.002)))
                continue
            ts = time.perf_counter()
            it, dn, rr = self._status()
            self.timing["status"] += time.perf_counter() - ts
            for e in active:
                # algorithmic bytes of the grouped products this chunk ran for the segment:
                # per nonzero its value and column (16-bit or int32; SELL3: one 16-bit
                # column per nine values), per row z (gathered once), p and q read and
                # written
                ran = int(it[e["seg"].slot]) - e.get("it_seen", 0)
                e["it_seen"] = int(it[e["seg"].slot])
                self.stats["iterations"] += ran
                if e["seg"].format == _lib.MIFE_FMT_SELL3:
                    cb = 2.0 / 9.0
                else:
                    cb = 2.0 if e["cbo_used"] else 4.0
                vb = 4.0 if e["mixed"] else 8.0
                self.stats["spmv_bytes"] += ran * ((vb + cb) * e["nnz"] + 5.0 * vb * e["nf"])
            finished = [e for e in active if dn[e["seg"].slot]]
            tr = time.perf_counter()
            if finished:
                tbl = self._build_table(finished, self.resid_table)
                call("mife_grouped_cg_residual", len(finished), tbl, self.max_segments, self.max_blocks,
                     C.byref(self.pools), ptr(self.b), ptr(self.x), ptr(self.resid_table), ptr(self.state),
                     fem._stream())
                it, dn, rr = self._status()
                redo = []
                for e in finished:
                    slot = e["seg"].slot
                    e["req"].iterations += int(it[slot])
                    e["req"].relres = float(rr[slot])
                    if rr[slot] > e["req"].rtol and e.get("restarts", 0) < 4:
                        # the recurrence drifted below the true residual: restart CG from it
                        e["restarts"] = e.get("restarts", 0) + 1
                        e["it_seen"] = 0
                        redo.append(e)
                    else:
                        self._retire(e, t0)
                if redo:
                    nbr = self._build_table(redo, self.resid_table)
                    self._init(len(redo), nbr, self.resid_table)
                redo_ids = {id(e) for e in redo}
                retired = {id(e) for e in finished if id(e) not in redo_ids}
                active = [e for e in active if id(e) not in retired]
                changed = True
            self.timing["retire"] += time.perf_counter() - tr
        for fut in self._pending.values():
            fut.result()
        self._pending.clear()
        return queue

    def _run_chunk(self, n_seg, nb):
        # a capture needs the structure worker idle (its allocations may synchronise
        # the device, which would invalidate the capture); only this thread submits
        # builds, so none starts during the capture
        worker_busy = any(not f.done() for f in self._pending.values())
        if not self.use_graph or (self.graph is None and worker_busy):
            self._iterate(n_seg, nb, self.chunk)
            return
        if self.graph is None:
            g = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    self._iterate(n_seg, nb, self.chunk)
            torch.cuda.current_stream().wait_stream(s)
            self.graph = g
            self.stats["graph_captures"] += 1
            # the capture did not run the iterations: replay once now
        self.graph.replay()
        self.stats["graph_replays"] += 1


def solve_one_fp64(req: Request, structures: dict, device=None) -> Request:
    """The one-at-a-time FP64 path for a request: assemble, reduce, Jacobi CG, field
    update (the single-solve engine on the request's own structures)."""
    dev = fem._dev(device)
    key = (req.kind, tuple(req.n_elem))
    st = structures.get(key)
    if st is None:
        st = _Structure(req.kind, req.n_elem, dev)
        structures.clear()
        structures[key] = st
    coef = st.coefficients(req.seed, dev)
    vals = torch.empty(st.nnz, dtype=torch.float64, device=dev)
    st.assemble(coef, vals, torch.empty(st.stage_bytes, dtype=torch.uint8, device=dev))
    red, S = st.red, st.S
    b = red.rhs(vals, st.u0, rhs=st.rhs)
    x = torch.zeros(red.n_free, dtype=torch.float64, device=dev)
    ws = fem.KrylovWorkspace(red.n_free, dev)
    info = fem.cg(S, S.gather(vals), b, x, req.rtol, 200000, ws)
    u = st.u0.clone()
    red.field_update(x, u)
    req.u = u.cpu().numpy()
    req.iterations = info.iterations
    req.relres = info.relres
    return req


def make_stream(n: int, seed: int = 2024, lo: int = 8, hi: int = 64, frac_diffusion: float = 0.7,
                rate: float | None = None):
    """n requests: 70% diffusion / 30% elasticity, N_x, N_y, N_z ~ U{lo..hi}, per-request
    coefficient seeds, Poisson arrivals at ``rate`` per second (all at t=0 if None)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n):
        kind = "diffusion" if rng.random() < frac_diffusion else "elasticity"
        ne = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        if rate:
            t += rng.exponential(1.0 / rate)
        reqs.append(Request(i, kind, ne, int(rng.integers(1 << 30)), arrival=t if rate else 0.0))
    return reqs
