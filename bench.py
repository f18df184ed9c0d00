#!/usr/bin/env python
"""Benchmark of iron's element-assembly + sparse-solve hot path on one MI355X.

Workload (default ``diffusion48``): 3-D Poisson-type diffusion on a 48x48x48
triquadratic Lagrange hex mesh (110,592 elements, 97^3 nodes, 3x3x3 Gauss),
per-element conductivity lognormal(0,1) from a fixed seed, source f = 1, u = 0 on
every boundary face.  One step is one full pass of the path:
  FP64 element integration written into the CSR rows (structured assembly) ->
  K_ff gather into FP64 and, rounded, FP32 solver values + RHS -> FP32 Jacobi CG
  inside FP64 iterative refinement to ||b - A x|| <= 1e-8 ||b|| -> field update ->
  back-substitution (reactions K u - f on every equations row).
The FP32 element-integration rate is measured apart from the pass.
The plain FP64 drop-in (FP64 assembly + FP64 Jacobi CG to 1e-8) is timed in the
same run, and the C restatement of iron's CPU algorithm on a bounded sample.

Further legs: the 16^3 trilinear Laplace, the 96x32x32 trilinear elasticity
cantilever, the 12^3 Q2/Q1 Mooney-Rivlin stretch, one x-slab of the 48^3 case
and the served stream of 1024 requests, each with its CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs main,mr,elastic,laplace,slab,serve]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="diffusion48", choices=["diffusion48", "laplace16"])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-seconds", type=float, default=16.0, help="budget of the main CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--format", default="sell", choices=["sell", "csr"], help="solver-matrix storage")
    ap.add_argument("--inner-rtol", type=float, default=0.1,
                    help="FP32 residual reduction between FP64 residual replacements")
    ap.add_argument("--no-mr", action="store_true", help="skip the Mooney-Rivlin (12^3 Q2/Q1) leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run three streaming passes y = 1.0 * x over 256 MiB FP32 arrays (reads and "
                         "writes 268,435,456 bytes each: the reference kernel for HBM byte counters)")
    ap.add_argument("--serve-requests", type=int, default=1024,
                    help="requests of the served-stream leg (0 skips it)")
    ap.add_argument("--serve-policies", default="auto,split,mixed,fp64",
                    help="precision policies of the served stream, each measured at capacity and at Poisson "
                         "arrivals")
    ap.add_argument("--legs", default="main,mr,elastic,laplace,slab,serve",
                    help="comma-separated legs to run: main (the workload above), mr, elastic (96x32x32 "
                         "elasticity), laplace (16^3 Laplace), slab (one eighth of the 48^3 case), serve")
    return ap.parse_args()


TRAFFIC_FILE = ROOT / "profiles" / "cg_spmv_diffusion48_q2.md"
WORKLOAD_TEXT = {
    "diffusion48": "diffusion48: 48^3 Q2 hex, 3x3x3 Gauss, kappa~lognormal(0,1) seed {seed}, f=1, u=0 on all "
                   "faces, FP32 inner CG + FP64 refinement to 1e-8, back-substitution",
    "diffusion48slab": "diffusion48 slab: 6x48x48 Q2 hex on [0,1/8]x[0,1]^2 (one of eight x-slabs of the 48^3 "
                       "case), 3x3x3 Gauss, kappa~lognormal(0,1) seed {seed}, f=1, u=0 on all faces, FP32 inner CG "
                       "+ FP64 refinement to 1e-8",
    "laplace16": "laplace16: 16^3 Q1 hex, 2x2x2 Gauss, kappa=1, u=0 at x=0 and u=1 at x=1, CG + Jacobi to 1e-10, "
                 "back-substitution",
}


def measured_traffic():
    """HBM bytes per launch of the fused FP32 SpMV on diffusion48, from the
    calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE summary kept with the profiles
    (line 'traffic_bytes_per_launch: N'), or None when that summary is absent."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def build_problem(args):
    import torch
    from iron_amd import basis, fem, mesh
    if args.workload == "diffusion48":
        ne, interp, extent, rtol = (48, 48, 48), basis.QUADRATIC_LAGRANGE, (1.0, 1.0, 1.0), 1e-8
    elif args.workload == "diffusion48slab":
        ne, interp, extent, rtol = (6, 48, 48), basis.QUADRATIC_LAGRANGE, (0.125, 1.0, 1.0), 1e-8
    elif args.workload == "laplace16":
        ne, interp, extent, rtol = (16, 16, 16), basis.LINEAR_LAGRANGE, (1.0, 1.0, 1.0), 1e-10
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    b = basis.basis(3, interp)
    m = mesh.RegularMesh(ne, [b], extent)
    dm = mesh.DofMap(m, [0])
    dev = torch.device("cuda:0")
    E = m.num_elements
    if args.workload in ("diffusion48", "diffusion48slab"):
        g = torch.Generator().manual_seed(args.seed)
        kappa = torch.randn(E, dtype=torch.float64, generator=g).exp()
        X = m.coords
        hi = np.array(extent)[None, :]
        fixed = np.any((np.abs(X) < 1e-12) | (np.abs(X - hi) < 1e-12), axis=1)
        ufix = np.zeros(dm.num_dofs)
        source = torch.ones(E, dtype=torch.float64)
    else:
        kappa = torch.ones(E, dtype=torch.float64)
        X = m.coords
        fixed = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1.0) < 1e-12)
        ufix = np.where(np.abs(X[:, 0] - 1.0) < 1e-12, 1.0, 0.0)
        source = None
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, dev, colours=mesh.structured_colours(m))
    asm = fem.ScalarAssembler(b, m.coords, m.elem_nodes[0], csr, grid=ne)
    red = fem.DirichletReduction(csr, fixed)
    S = fem.SolverMatrix(red, args.format)
    P = dict(m=m, dm=dm, basis=b, csr=csr, asm=asm, red=red, S=S, rtol=rtol, fixed=fixed, ufix=ufix,
             kappa=kappa.to(dev), source=None if source is None else source.to(dev), kappa_h=kappa.numpy(),
             ws=fem.KrylovWorkspace(red.n_free, dev))
    P["v64"] = csr.zeros(fem.F64)
    P["v32"] = csr.zeros(fem.F32)
    P["src"] = torch.zeros(dm.num_dofs, dtype=torch.float64, device=dev)
    P["u"] = torch.as_tensor(np.where(fixed, ufix, 0.0), device=dev)
    P["u0"] = P["u"].clone()
    P["R"] = torch.empty(dm.num_dofs, dtype=torch.float64, device=dev)
    P["sv64"] = torch.empty(max(S.stored, 1), dtype=torch.float64, device=dev)
    P["sv32"] = torch.empty(max(S.stored, 1), dtype=torch.float32, device=dev)
    P["b"] = torch.empty(red.n_free, dtype=torch.float64, device=dev)
    P["x"] = torch.zeros(red.n_free, dtype=torch.float64, device=dev)
    return P


def step(P, mixed: bool, inner_rtol: float):
    """One pass: FP64 assembly, solver values and RHS, the solve, field update and
    the back-substitution of the reactions.  The mixed solve's FP32 values are the
    FP64 solver values rounded, written by the same gather."""
    from iron_amd import fem
    v64, src = P["v64"], P["src"]
    v64.zero_()
    src.zero_()
    P["asm"].assemble(v64, kappa=P["kappa"], source=P["source"], src_vec=src if P["source"] is not None else None,
                      prec=fem.F64)
    red, S = P["red"], P["S"]
    if mixed:
        S.gather2(v64, out64=P["sv64"], out32=P["sv32"])
    else:
        S.gather(v64, fem.F64, out=P["sv64"])
    P["u"].copy_(P["u0"])
    red.rhs(v64, P["u"], src=src if P["source"] is not None else None, out=P["b"])
    x = P["x"]
    x.zero_()
    if mixed:
        info = fem.refine(S, P["sv64"], P["sv32"], P["b"], x, P["rtol"], 200000, P["ws"], inner_rtol=inner_rtol)
    else:
        info = fem.cg(S, P["sv64"], P["b"], x, P["rtol"], 200000, P["ws"], prec=fem.F64)
    red.field_update(x, P["u"])
    fem.backsubstitute(P["csr"], v64, P["u"], src=src if P["source"] is not None else None, out=P["R"])
    return info


def fp32_assembly_profile(P, runs=3):
    """Kernel profile of `runs` FP32 assemblies of the problem (the FP32 element-matrix
    rate, measured apart from the pass, which rounds the FP64 values instead)."""
    import torch
    from iron_amd import _lib, fem
    v32 = P["v32"]
    _lib.profile_enable(True)
    _lib.profile_reset()
    for _ in range(runs):
        v32.zero_()
        P["asm"].assemble(v32, kappa=P["kappa"], prec=fem.F32)
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    return prof


def timed_steps(P, k, mixed, inner_rtol):
    import torch
    lat, infos = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        t = time.perf_counter()
        infos.append(step(P, mixed, inner_rtol))
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, lat, infos


# ------------------------------------------------------------------ CPU baselines
def host_threads():
    """Host threads for the all-cores CPU baseline: OMP_NUM_THREADS when set (the
    host-core share given to this GPU), else the CPUs this process may run on."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def oracle_lib():
    """The C restatement of iron's CPU path (oracle/liboracle_fe.so), or None."""
    so = ROOT / "oracle" / "liboracle_fe.so"
    if not so.exists():
        return None
    lib = C.CDLL(str(so))
    for name in ("oracle_assemble_scalar3", "oracle_assemble_scalar3_par", "oracle_assemble_elastic3",
                 "oracle_assemble_elastic3_par", "oracle_mr_assemble"):
        getattr(lib, name).restype = C.c_int
    lib.oracle_reduce.restype = C.c_int64
    lib.oracle_cg.restype = C.c_int
    lib.oracle_cg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                              C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.oracle_gmres.restype = C.c_int
    lib.oracle_gmres.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    return lib


def _sampled_assembly(assemble, E, budget_s, threads):
    """Times assemble(e0, e1, nt) over elements [0, Es), Es sized to the budget;
    completes the matrix untimed; returns the full-assembly estimate and Es."""
    nthr = host_threads()
    t = time.perf_counter()
    probe = min(E, 256 * threads)
    assemble(0, probe, threads)
    per_elem = (time.perf_counter() - t) / probe
    Es = int(min(E, max(probe, budget_s / max(per_elem, 1e-9))))
    return per_elem, Es, nthr


def _cg_sample(lib, n, srp, scol, sv, bb, iters_needed, budget_s, threads):
    """Per-iteration time of the C Jacobi CG on the real system, over a bounded
    number of iterations (at most iters_needed)."""
    x = np.zeros(n)
    rr = C.c_double(0.0)
    t = time.perf_counter()
    it_probe = 5
    lib.oracle_cg(n, _P(srp), _P(scol), _P(sv), _P(bb), _P(x), 0.0, it_probe, threads, C.byref(rr))
    per_it = (time.perf_counter() - t) / it_probe
    n_it = int(max(it_probe, min(iters_needed, budget_s / max(per_it, 1e-9))))
    x[:] = 0.0
    t = time.perf_counter()
    lib.oracle_cg(n, _P(srp), _P(scol), _P(sv), _P(bb), _P(x), 0.0, n_it, threads, C.byref(rr))
    return (time.perf_counter() - t) / n_it, n_it


def _reduce_and_backsub(lib, csr_rp, csr_col, vals, new_idx, u, src, n_free, threads):
    """The C reduction (K_ff + lifted RHS, serial as iron's per-rank loop) and the
    back-substitution over every equations row: (srp, scol, sv, b, t_reduce, t_backsub)."""
    n = len(csr_rp) - 1
    srp = np.zeros(n_free + 1, dtype=np.int32)
    nnz = int(csr_rp[-1])
    scol = np.zeros(max(nnz, 1), dtype=np.int32)
    sv = np.zeros(max(nnz, 1))
    bb = np.zeros(n_free)
    t = time.perf_counter()
    lib.oracle_reduce(n, _P(csr_rp), _P(csr_col), _P(vals), _P(new_idx), _P(u), _P(src), _P(srp), _P(scol), _P(sv),
                      _P(bb))
    t_red = time.perf_counter() - t
    R = np.zeros(n)
    t = time.perf_counter()
    lib.oracle_backsubstitute(n, _P(csr_rp), _P(csr_col), _P(vals), _P(u), _P(src), _P(R), threads)
    t_bs = time.perf_counter() - t
    return srp, scol, sv, bb, t_red, t_bs


def cpu_baseline_scalar(P, fp64_iters, budget_s, threads=1):
    """The C restatement of iron's FP64 path for a scalar problem, timed over a
    bounded sample and scaled to one full pass: element loop with bisection CSR add
    over a sample of elements (scaled to all), the K_ff/RHS reduction and the
    back-substitution in full, and a bounded number of Jacobi-CG iterations scaled to
    the iterations the FP64 solve needs.  threads > 1: elements shared out with
    atomic adds, CG rows and dot products split (a proxy for iron on that many ranks)."""
    lib = oracle_lib()
    if lib is None:
        return None
    m, b, csr, red = P["m"], P["basis"], P["csr"], P["red"]
    E = m.num_elements
    rp = csr.row_ptr.cpu().numpy().astype(np.int32)
    col = csr.col[: csr.nnz].cpu().numpy().astype(np.int32)
    ed = P["dm"].element_dofs()
    en = np.ascontiguousarray(m.elem_nodes[0], dtype=np.int32)
    coords = np.ascontiguousarray(m.coords, dtype=np.float64)
    dN = np.ascontiguousarray(b.dN, dtype=np.float64)
    N = np.ascontiguousarray(b.N, dtype=np.float64)
    w = np.ascontiguousarray(b.weights, dtype=np.float64)
    kap = np.ascontiguousarray(P["kappa_h"], dtype=np.float64)
    f = np.ones(E) if P["source"] is not None else None
    vals = np.zeros(csr.nnz)
    src = np.zeros(csr.n)

    def assemble(e0, e1, nt):
        args = (E, b.num_nodes, b.num_gauss, _P(dN), _P(N), _P(w), _P(coords), _P(en), _P(ed), _P(kap), _P(f),
                _P(rp), _P(col), _P(vals), _P(src), e0, e1)
        return lib.oracle_assemble_scalar3(*args) if nt == 1 else lib.oracle_assemble_scalar3_par(*args, nt)

    per_elem, Es, nthr = _sampled_assembly(assemble, E, 0.4 * budget_s, threads)
    vals[:] = 0.0
    src[:] = 0.0
    t = time.perf_counter()
    assemble(0, Es, threads)
    t_asm = (time.perf_counter() - t) * (E / Es)
    if Es < E:   # complete the matrix (untimed, all host threads): the solve runs on the real system
        assemble(Es, E, nthr)
    new_idx = red.new_idx.cpu().numpy().astype(np.int32)
    u = P["u0"].cpu().numpy()
    srp, scol, sv, bb, t_red, t_bs = _reduce_and_backsub(lib, rp, col, vals, new_idx, u, src if f is not None else None,
                                                         red.n_free, threads)
    per_it, n_it = _cg_sample(lib, red.n_free, srp, scol, sv, bb, fp64_iters, 0.5 * budget_s, threads)
    t_solve = per_it * fp64_iters
    t_total = t_asm + t_red + t_solve + t_bs
    return {
        "value": E / t_total, "unit": "elems/s", "cores": threads, "kind": "port",
        "sample": (f"C restatement of iron's FP64 path on {threads} thread(s) ({cpu_model()}): element loop with "
                   f"bisection CSR add over {Es} of {E} elements (x{E / Es:.2f}), K_ff/RHS reduction (serial) and "
                   f"back-substitution in full, {n_it} Jacobi-CG iterations scaled to the {fp64_iters} iterations "
                   f"the FP64 solve needs"),
        "assembly_s": t_asm, "reduce_s": t_red, "backsub_s": t_bs, "cg_iter_ms": per_it * 1e3, "solve_s": t_solve,
        "pass_s": t_total, "solves_per_s": 1.0 / t_total,
    }


def cpu_baseline_elastic(E_count, b, coords, en, ed, mat, csr_rp, csr_col, new_idx, n_free, rhs, fp64_iters,
                         budget_s, threads=1):
    """As cpu_baseline_scalar for 3-D linear elasticity (iron's element loop with its
    DXI_DX index order, bisection add), the reduction with the traction RHS, the
    back-substitution and Jacobi CG."""
    lib = oracle_lib()
    if lib is None:
        return None
    dN = np.ascontiguousarray(b.dN, dtype=np.float64)
    w = np.ascontiguousarray(b.weights, dtype=np.float64)
    vals = np.zeros(int(csr_rp[-1]))

    def assemble(e0, e1, nt):
        args = (E_count, b.num_nodes, b.num_gauss, _P(dN), _P(w), _P(coords), _P(en), _P(ed), _P(mat), _P(csr_rp),
                _P(csr_col), _P(vals), e0, e1)
        return lib.oracle_assemble_elastic3(*args) if nt == 1 else lib.oracle_assemble_elastic3_par(*args, nt)

    per_elem, Es, nthr = _sampled_assembly(assemble, E_count, 0.4 * budget_s, threads)
    vals[:] = 0.0
    t = time.perf_counter()
    assemble(0, Es, threads)
    t_asm = (time.perf_counter() - t) * (E_count / Es)
    if Es < E_count:
        assemble(Es, E_count, nthr)
    u = np.zeros(len(csr_rp) - 1)
    srp, scol, sv, bb, t_red, t_bs = _reduce_and_backsub(lib, csr_rp, csr_col, vals, new_idx, u, rhs, n_free,
                                                         threads)
    per_it, n_it = _cg_sample(lib, n_free, srp, scol, sv, bb, fp64_iters, 0.5 * budget_s, threads)
    t_solve = per_it * fp64_iters
    t_total = t_asm + t_red + t_solve + t_bs
    return {
        "value": E_count / t_total, "unit": "elems/s", "cores": threads, "kind": "port",
        "sample": (f"C restatement of iron's FP64 linear-elasticity path on {threads} thread(s) ({cpu_model()}): "
                   f"element loop with bisection CSR add over {Es} of {E_count} elements (x{E_count / Es:.2f}), "
                   f"reduction (serial) and back-substitution in full, {n_it} Jacobi-CG iterations scaled to the "
                   f"{fp64_iters} iterations the FP64 solve needs"),
        "assembly_s": t_asm, "reduce_s": t_red, "backsub_s": t_bs, "cg_iter_ms": per_it * 1e3, "solve_s": t_solve,
        "pass_s": t_total, "solves_per_s": 1.0 / t_total,
    }


# ------------------------------------------------------------------ Mooney-Rivlin
def mr_flops_per_element(nn, npd, ng):
    """Algorithmic FP64 flops of one Mooney-Rivlin element (residual + exact Jacobian),
    counted from the kernel's loops: per Gauss point 72 nn^2 (displacement block:
    (3nn)^2 entries x 3 FMAs + weight), 54 nn npd (pressure coupling), 216 nn (metrics,
    reference gradients, T = GX . A) and about 4,200 for kinematics, stress and the
    81-entry tangent."""
    return ng * (72 * nn * nn + 54 * nn * npd + 216 * nn + 4200)


def cpu_baseline_mr(sysg, newton, x_final, mat, jac_assemblies, res_assemblies, gmres_iters, restart, budget_s):
    """The C restatement of the Newton path on one core: the Mooney-Rivlin element
    loop (residual + exact Jacobian, bisection add) over a sample of elements scaled
    to the whole mesh, the residual-only loop likewise, and GMRES(restart) with the
    pressure-block diagonal preconditioner on the real Newton matrix at the final
    state over a bounded number of iterations; scaled to the Jacobian and residual
    assemblies and the GMRES iterations the GPU run needed."""
    lib = oracle_lib()
    if lib is None:
        return None
    rm = sysg.rm
    bu, bp = sysg.bu, sysg.bp
    E, nn = rm.num_elements, bu.num_nodes
    dN = np.ascontiguousarray(bu.dN, dtype=np.float64)
    w = np.ascontiguousarray(bu.weights, dtype=np.float64)
    psi = np.ones((bu.num_gauss, 1)) if bp is None else np.ascontiguousarray(bp.N, dtype=np.float64)
    npd = psi.shape[1]
    coords = np.ascontiguousarray(sysg.X)
    state = np.ascontiguousarray(x_final.cpu().numpy())
    en = np.ascontiguousarray(rm.elem_nodes[0], dtype=np.int32)
    ed = np.ascontiguousarray(sysg.elem_dofs, dtype=np.int32)
    matc = np.ascontiguousarray(mat, dtype=np.float64)
    rp, col = sysg.csr.to_host()
    rp, col = rp.astype(np.int32), col.astype(np.int32)
    vals = np.zeros(len(col))
    R = np.zeros(sysg.ndofs)

    def assemble(e0, e1, with_jac):
        return lib.oracle_mr_assemble(E, nn, npd, bu.num_gauss, sysg.nu, _P(dN), _P(psi), _P(w), _P(coords),
                                      _P(state), _P(en), _P(ed), _P(matc), _P(rp), _P(col),
                                      _P(vals) if with_jac else None, _P(R), e0, e1)

    t = time.perf_counter()
    assemble(0, 8, True)
    per = (time.perf_counter() - t) / 8
    Es = int(min(E, max(8, 0.3 * budget_s / max(per, 1e-9))))
    vals[:] = 0.0
    R[:] = 0.0
    t = time.perf_counter()
    assemble(0, Es, True)
    t_jac = (time.perf_counter() - t) * E / Es
    t = time.perf_counter()
    assemble(0, Es, False)
    t_res = (time.perf_counter() - t) * E / Es
    vals[:] = 0.0
    R[:] = 0.0
    assemble(0, E, True)
    new_idx = newton.red.new_idx.cpu().numpy().astype(np.int32)
    n_free = newton.red.n_free
    srp, scol, sv, bb, _, _ = _reduce_and_backsub(lib, rp, col, vals, new_idx, np.zeros(sysg.ndofs), None, n_free, 1)
    bb = -R[newton.red.free_idx.cpu().numpy()]
    pc = np.ascontiguousarray(newton.pc_neg.cpu().numpy())
    xg = np.zeros(n_free)
    rr = C.c_double(0.0)
    t = time.perf_counter()
    lib.oracle_gmres(n_free, _P(srp), _P(scol), _P(sv), _P(pc), _P(bb), _P(xg), 0.0, restart, 8, 1, C.byref(rr))
    per_it = (time.perf_counter() - t) / 8
    n_it = int(max(restart, min(gmres_iters, 0.5 * budget_s / max(per_it, 1e-9))))
    n_it = (n_it // restart) * restart if n_it >= restart else n_it
    xg[:] = 0.0
    t = time.perf_counter()
    lib.oracle_gmres(n_free, _P(srp), _P(scol), _P(sv), _P(pc), _P(bb), _P(xg), 0.0, restart, n_it, 1, C.byref(rr))
    per_it = (time.perf_counter() - t) / n_it
    total = jac_assemblies * t_jac + res_assemblies * t_res + gmres_iters * per_it
    return {
        "value": 1.0 / total, "unit": "solves/s (10 load increments)", "cores": 1, "kind": "port",
        "sample": (f"C restatement on 1 thread ({cpu_model()}): Mooney-Rivlin residual + exact Jacobian element "
                   f"loop with bisection add over {Es} of {E} elements (x{E / Es:.2f}), the residual-only loop "
                   f"likewise, and {n_it} GMRES({restart}) iterations on the final Newton matrix; scaled to the "
                   f"{jac_assemblies} Jacobian and {res_assemblies} residual assemblies and {gmres_iters} GMRES "
                   f"iterations of the GPU run"),
        "jacobian_assembly_s": t_jac, "residual_assembly_s": t_res, "gmres_iter_ms": per_it * 1e3,
        "solve_s": total, "assembly_elems_per_s": E / t_jac,
    }


def mr_leg(runs, cpu_seconds):
    """12^3 triquadratic/trilinear Mooney-Rivlin cube (c1=2, c2=6), rollers on x=0, y=0,
    z=0, x=1 face stretched to 1.2 in 10 load increments, Newton to a relative residual
    of 1e-8 per increment: the default preconditioned-MINRES steps with Eisenstat-Walker
    tolerances, and GMRES(256) steps each to 1e-8 as the configuration states it."""
    import torch
    from iron_amd import _lib, basis, finite_elasticity as fe, mesh
    rm = mesh.RegularMesh((12, 12, 12), [basis.basis(3, 2, 3), basis.basis(3, 1, 3)], (1.0, 1.0, 1.0))
    sysg = fe.MRSystem(rm, 1, fe.DEVIATORIC)
    nu, X = sysg.nu, rm.coords
    fixed = np.zeros(sysg.ndofs, bool)
    x0 = np.concatenate([X[:, 0], X[:, 1], X[:, 2], np.zeros(sysg.npd)])
    target = x0.copy()
    for c in range(3):
        fixed[c * nu:(c + 1) * nu] |= X[:, c] == 0.0
    fixed[:nu] |= X[:, 0] == 1.0
    target[:nu][X[:, 0] == 1.0] = 1.2
    mat = np.tile([2.0, 6.0], (rm.num_elements, 1))
    newton = fe.MRNewton(sysg, fixed, mat)
    xe = X * np.array([1.2, 1.2 ** -0.5, 1.2 ** -0.5])
    E = rm.num_elements
    fl = mr_flops_per_element(27, 8, 27) * E

    def run(opts):
        x = torch.as_tensor(x0, device=sysg.device)
        infos = fe.load_increments(newton, x, 10, fixed, target, opts=opts)
        torch.cuda.synchronize()
        return x, infos

    def profiled(opts, reps):
        """reps timed solves with the library's timing events off, then one more with
        them on for the per-kernel times (the events cost the solve ~25% on GMRES)."""
        run(opts)
        times = []
        for _ in range(reps):
            t = time.perf_counter()
            x, infos = run(opts)
            times.append(time.perf_counter() - t)
        _lib.profile_enable(True)
        _lib.profile_reset()
        run(opts)
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        return x, infos, times, prof

    opts = fe.NewtonOptions(atol=1e-12, rtol=1e-8, stol=0.0, linear_rtol="ew")
    x, infos, times, prof = profiled(opts, runs)
    err = float(np.abs(x.cpu().numpy()[:3 * nu].reshape(3, -1).T - xe).max())
    n_j, ms_j = prof.get("finite_elasticity", (0, 0.0))
    n_r, ms_r = prof.get("finite_elasticity_residual", (0, 0.0))
    avg_j = ms_j / max(n_j, 1)
    avg_r = ms_r / max(n_r, 1)
    ach = fl / (avg_j * 1e-3) / 1e12 if avg_j > 0 else None
    out = {
        "workload": "mr12: 12^3 Q2/Q1 Mooney-Rivlin (deviatoric, c1=2, c2=6), rollers + x-stretch to 1.2 in 10 "
                    "increments, Newton to 1e-8 per increment, MINRES steps with Eisenstat-Walker tolerances",
        "elements": E, "dofs": sysg.ndofs, "free_dofs": newton.red.n_free, "jacobian_nnz": sysg.csr.nnz,
        "solve_s": float(np.median(times)), "solves_per_s": 1.0 / float(np.median(times)),
        "newton_iterations": sum(i.iterations for i in infos),
        "minres_iterations": sum(i.linear_iterations for i in infos),
        "max_node_error_vs_closed_form": err,
        "jacobian_assembly_ms": avg_j, "jacobian_assemblies": n_j // max(runs, 1),
        "residual_assembly_ms": avg_r, "residual_assemblies": n_r // max(runs, 1),
        "assembly_elems_per_s": E / (avg_j * 1e-3) if avg_j > 0 else None,
        "roofline": {"kernel": "mr_element_kernel<27,8,27> (residual + exact Jacobian, FP64), residual+Jacobian "
                               "launches only", "bound": "fp64", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS if ach else None,
                     "flops_per_element": mr_flops_per_element(27, 8, 27), "avg_launch_ms": avg_j,
                     "launches": n_j},
        "kernel_ms": {k: v for k, v in prof.items() if v[0]},
    }
    # the configuration's solver: GMRES, pressure-block preconditioner, each Newton step
    # to 1e-8; restart 256 (the longest the library takes: on these saddle-point
    # systems GMRES(100) needs about 2.5x the iterations)
    restart = 256
    opts_g = fe.NewtonOptions(atol=1e-12, rtol=1e-8, stol=0.0, linear_rtol=1e-8, linear="gmres",
                              gmres_restart=restart)
    xg, infos_g, times_g, prof_g = profiled(opts_g, 1)
    gi = sum(i.linear_iterations for i in infos_g)
    out["gmres_1e8"] = {"solve_s": times_g[0], "newton_iterations": sum(i.iterations for i in infos_g),
                        "gmres_iterations": gi, "restart": restart,
                        "us_per_gmres_iteration_incl_newton": times_g[0] / max(gi, 1) * 1e6,
                        "max_node_error_vs_closed_form":
                            float(np.abs(xg.cpu().numpy()[:3 * nu].reshape(3, -1).T - xe).max()),
                        "kernel_ms": {k: v for k, v in prof_g.items() if v[0]}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline_mr(sysg, newton, xg, mat, prof_g.get("finite_elasticity", (0, 0))[0],
                                              prof_g.get("finite_elasticity_residual", (0, 0))[0], gi, restart,
                                              cpu_seconds)
    return out


# ------------------------------------------------------------------ elasticity
def elastic_leg(runs=3, cpu_seconds=10.0):
    """96x32x32 trilinear linear elasticity on [0,3]x[0,1]^2: lognormal Young's modulus
    10 exp(0.3 N(0,1)) per element (seed 7), Poisson ratio 0.3, clamp at x = 0,
    uniform traction t_z = -1 on the x = 3 face (NEUMANN_POINT -1 on its nodes, the side
    faces next to it left out by NEUMANN_INTEGRATED_ONLY DOFs one element in, so the
    nodal forces sum to -1 x the face area).  One pass: FP64 (and in the mixed pass FP32
    MFMA) element stiffness, solver values and RHS, Jacobi CG to a 1e-8 relative
    residual (SELL3: 3x3 node blocks, 16-bit node columns), field update, back-substitution."""
    import torch
    from iron_amd import _lib, basis, engine, fem, mesh
    ne = (96, 32, 32)
    b = basis.basis(3, basis.LINEAR_LAGRANGE)
    m = mesh.RegularMesh(ne, [b], (3.0, 1.0, 1.0))
    dm = mesh.DofMap(m, [0, 0, 0])
    dev = torch.device("cuda:0")
    X = m.coords
    E, N = m.num_elements, m.num_nodes(0)
    g = torch.Generator().manual_seed(7)
    Ee = 10.0 * torch.exp(0.3 * torch.randn(E, dtype=torch.float64, generator=g)).numpy()
    mat_h = np.ascontiguousarray(np.stack([Ee, Ee, Ee, np.full(E, 0.3), np.full(E, 0.3), np.full(E, 0.3)], axis=1))
    mat = torch.as_tensor(mat_h, device=dev)
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, dev, colours=mesh.structured_colours(m))
    asm = fem.ElasticAssembler(b, X, m.elem_nodes[0], csr, grid=ne)
    clamp = X[:, 0] == 0.0
    red = fem.DirichletReduction(csr, np.concatenate([clamp, clamp, clamp]))
    S = fem.SolverMatrix(red, interleave=(3, N), block=True)
    tip = np.nonzero(X[:, 0] == 3.0)[0]
    hx = 3.0 / ne[0]
    onb = (X[:, 1] == 0.0) | (X[:, 1] == 1.0) | (X[:, 2] == 0.0) | (X[:, 2] == 1.0)
    ring = frozenset(int(2 * N + i) for i in np.nonzero(np.isclose(X[:, 0], 3.0 - hx) & onb)[0])
    ft_h = engine.neumann_integrate(m, dm, [b] * 3, {int(2 * N + i): -1.0 for i in tip}, integrated_only=ring)
    ft = torch.as_tensor(ft_h, device=dev)
    v, v32 = csr.zeros(fem.F64), csr.zeros(fem.F32)
    sv64 = torch.empty(max(S.stored, 1), dtype=torch.float64, device=dev)
    sv32 = torch.empty(max(S.stored, 1), dtype=torch.float32, device=dev)
    u = torch.zeros(csr.n, dtype=torch.float64, device=dev)
    R = torch.empty(csr.n, dtype=torch.float64, device=dev)
    bvec = torch.empty(red.n_free, dtype=torch.float64, device=dev)
    x = torch.zeros(red.n_free, dtype=torch.float64, device=dev)
    ws = fem.KrylovWorkspace(red.n_free, dev)

    def one(mixed=True):
        v.zero_()
        asm.assemble(v, mat)
        if not mixed:
            S.gather(v, fem.F64, out=sv64)
        u.zero_()
        red.rhs(v, u, rhs=ft, out=bvec)
        x.zero_()
        if mixed:
            S.gather2(v, out64=sv64, out32=sv32)
            info = fem.refine(S, sv64, sv32, bvec, x, 1e-8, 200000, ws)
        else:
            info = fem.cg(S, sv64, bvec, x, 1e-8, 200000, ws, prec=fem.F64)
        red.field_update(x, u)
        fem.backsubstitute(csr, v, u, out=R)
        return info

    res = {}
    for mixed in (True, False):
        one(mixed)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(runs):
            info = one(mixed)
        torch.cuda.synchronize()
        res[mixed] = ((time.perf_counter() - t) / runs, info)
    _lib.profile_enable(True)
    _lib.profile_reset()
    one(True)
    v32.zero_()
    asm.assemble(v32, mat, prec=fem.F32)   # the FP32 element-matrix rate (the pass rounds FP64 values)
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    n_s, ms_s = prof.get("cg_spmv_f32", (0, 0.0))
    per = ms_s / max(n_s, 1)
    if S.block == 3:
        # SELL3: an FP32 value per nonzero and one 16-bit node column per 3x3 block
        alg = 4.0 * red.nnz + 2.0 * red.nnz / 9.0 + 24.0 * red.n_free
        kern = ("cg_spmv<float> (fused SpMV + p/q update + p.q) on SELL3 3x3 node blocks; bytes: 4 per nonzero, "
                "2 per block (16-bit node column), 24 per row")
    else:
        alg = 8.0 * red.nnz + 24.0 * red.n_free   # values + int32 columns, z, p/q read, p/q written
        kern = "cg_spmv<float> (fused SpMV + p/q update + p.q, 16-bit columns)"
    t_mx, info_mx = res[True]
    t_64, info_64 = res[False]
    a64 = prof.get("assemble_elastic", (0, 0.0))[1]
    a32 = prof.get("assemble_elastic_f32", (0, 0.0))[1]
    pick = fem.auto_precision(3, S.stored)
    t_pick = t_mx if pick == "mixed" else t_64
    out = {
        "workload": "elasticity96: 96x32x32 Q1 hex on [0,3]x[0,1]^2, 2x2x2 Gauss, E = 10 exp(0.3 N(0,1)) per element "
                    "(seed 7), nu = 0.3, clamp at x=0, uniform traction t_z=-1 on x=3 (nodal forces sum to -1), "
                    "Jacobi CG to 1e-8, back-substitution",
        "elements": E, "free_dofs": red.n_free, "solver_nnz": red.nnz, "sell_padding": S.padding,
        "applied_force_z": float(ft_h.sum()),
        "api_precision": pick, "api_ms_per_pass": t_pick * 1e3, "api_elems_per_s": E / t_pick,
        "ms_per_pass": t_mx * 1e3, "elems_per_s": E / t_mx, "iterations_inner": info_mx.iterations,
        "refinement_steps": info_mx.outer, "relres": info_mx.relres,
        "fp64_ms_per_pass": t_64 * 1e3, "fp64_iterations": info_64.iterations, "fp64_relres": info_64.relres,
        "assembly_ms_f64": a64, "assembly_ms_f32": a32,
        "assembly_elems_per_s_f32": E / (a32 * 1e-3) if a32 > 0 else None,
        "solver_format": S.fmt,
        "roofline": {"kernel": kern, "bound": "hbm",
                     "achieved": alg / (per * 1e-3) / 1e9 if per > 0 else 0.0, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (alg / (per * 1e-3) / 1e9 / HBM_PEAK_GBS) if per > 0 else 0.0,
                     "bytes_per_launch": alg, "avg_launch_ms": per, "launches": n_s},
    }
    if cpu_seconds > 0:
        rp = csr.row_ptr.cpu().numpy().astype(np.int32)
        col = csr.col[: csr.nnz].cpu().numpy().astype(np.int32)
        ed = np.ascontiguousarray(dm.element_dofs(), dtype=np.int32)
        en = np.ascontiguousarray(m.elem_nodes[0], dtype=np.int32)
        new_idx = red.new_idx.cpu().numpy().astype(np.int32)
        args = (E, b, np.ascontiguousarray(X, dtype=np.float64), en, ed, mat_h, rp, col, new_idx, red.n_free, ft_h,
                info_64.iterations)
        out["cpu_baseline"] = cpu_baseline_elastic(*args, cpu_seconds, 1)
        nthr = host_threads()
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline_elastic(*args, 0.5 * cpu_seconds, nthr)
    return out


# ------------------------------------------------------------------ Laplace
def laplace_leg(args, runs=5, cpu_seconds=4.0):
    """16^3 trilinear Laplace, u = 0 at x = 0 and u = 1 at x = 1: the full pass (FP64
    assembly, solver values, CG to 1e-10 in FP64 or FP32 CG with FP64 refinement, field
    update, back-substitution) in both precisions, and the one the API picks."""
    import argparse
    import torch
    from iron_amd import fem
    a = argparse.Namespace(**vars(args))
    a.workload = "laplace16"
    P = build_problem(a)
    out = {"workload": WORKLOAD_TEXT["laplace16"], "elements": P["m"].num_elements, "free_dofs": P["red"].n_free}
    for mixed, key in ((True, "mixed"), (False, "fp64")):
        timed_steps(P, 1, mixed, 0.1)
        wall, lat, infos = timed_steps(P, runs, mixed, 0.1)
        out[f"{key}_ms_per_pass"] = wall / runs * 1e3
        out[f"{key}_elems_per_s"] = P["m"].num_elements * runs / wall
        out[f"{key}_iterations"] = infos[-1].iterations
        out[f"{key}_relres"] = infos[-1].relres
    pick = fem.auto_precision(1, P["S"].stored)
    out["api_precision"] = pick
    out["api_ms_per_pass"] = out[f"{pick}_ms_per_pass"]
    torch.cuda.synchronize()
    if cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline_scalar(P, out["fp64_iterations"], cpu_seconds, 1)
        nthr = host_threads()
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline_scalar(P, out["fp64_iterations"], 0.5 * cpu_seconds, nthr)
    return out


def slab_leg(args, runs=3):
    """The 48^3 triquadratic diffusion case cut into eight slabs along x: one slab's
    kernels at their real per-GPU shape (6x48x48 elements; here with its own Dirichlet
    faces, no halo).  Reports the assembly and fused FP32 SpMV rates at that size."""
    import argparse
    import torch
    from iron_amd import _lib
    a = argparse.Namespace(**vars(args))
    a.workload = "diffusion48slab"
    P = build_problem(a)
    timed_steps(P, 1, True, args.inner_rtol)
    _lib.profile_enable(True)
    _lib.profile_reset()
    wall, lat, infos = timed_steps(P, runs, True, args.inner_rtol)
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    prof32 = fp32_assembly_profile(P)
    red = P["red"]
    n_s, ms_s = prof.get("cg_spmv_f32", (0, 0.0))
    per = ms_s / max(n_s, 1)
    alg = 8.0 * red.nnz + 24.0 * red.n_free
    E = P["m"].num_elements
    a64 = prof.get("assemble_scalar", (0, 0.0))
    a32 = prof32.get("assemble_scalar_f32", (0, 0.0))
    torch.cuda.synchronize()
    return {
        "workload": WORKLOAD_TEXT["diffusion48slab"].format(seed=args.seed),
        "elements": E, "free_dofs": red.n_free, "solver_nnz": red.nnz,
        "ms_per_pass": wall / runs * 1e3, "iterations_inner": infos[-1].iterations,
        "assembly_elems_per_s_f64": E / (a64[1] / max(a64[0], 1) * 1e-3) if a64[0] else None,
        "assembly_elems_per_s_f32": E / (a32[1] / max(a32[0], 1) * 1e-3) if a32[0] else None,
        "spmv_f32": {"avg_launch_ms": per, "bytes_per_launch": alg,
                     "achieved_GBs": alg / (per * 1e-3) / 1e9 if per > 0 else 0.0,
                     "frac": alg / (per * 1e-3) / 1e9 / HBM_PEAK_GBS if per > 0 else 0.0},
    }


# ------------------------------------------------------------------ served stream
def serve_leg(n, policies, seed=2024):
    """Served stream: n requests, 70% diffusion / 30% elasticity trilinear boxes with
    N_x, N_y, N_z ~ U{8..64}, lognormal(0,1) coefficients (per-request seeds), BCs of
    the 16^3 Laplace and 96x32x32 elasticity cases, each solved to a 1e-8 FP64 residual.
    For each precision policy: the grouped server's capacity (all requests at t=0) and
    latency (arrival -> solution on the host) at Poisson arrivals of 0.7x that capacity;
    then the one-at-a-time FP64 path on the same requests, and the grouped product's
    roofline on the first 256 requests."""
    import torch
    from iron_amd import _lib, serving
    pol_out = {}
    for pol in policies:
        srv = serving.Server(hbm_budget=60e9, max_segments=256, chunk=16, use_graph=True, precision=pol)
        srv.run(serving.make_stream(8, seed=seed + 1, lo=8, hi=16))
        for k in srv.timing:
            srv.timing[k] = 0.0
        reqs = serving.make_stream(n, seed=seed)
        t = time.perf_counter()
        out = srv.run(reqs)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        cap = n / wall
        timing = dict(srv.timing)
        for k in srv.timing:
            srv.timing[k] = 0.0
        reqs_p = serving.make_stream(n, seed=seed, rate=0.7 * cap)
        t = time.perf_counter()
        out_p = srv.run(reqs_p)
        torch.cuda.synchronize()
        wall_p = time.perf_counter() - t
        lat = np.array([r.latency for r in out_p]) * 1e3
        # structure waits happen inside the admission window (counted once, in admit)
        host = timing["admit"] + timing["retire"] + timing["chunk"]
        progress(f"served stream, policy {pol}: {cap:.1f} solves/s at capacity")
        pol_out[pol] = {
            "capacity_solves_per_s": cap, "elems_per_s_at_capacity": sum(int(np.prod(r.n_elem)) for r in out) / wall,
            "poisson_rate_per_s": 0.7 * cap, "poisson_solves_per_s": n / wall_p,
            "p50_latency_ms": float(np.percentile(lat, 50)), "p99_latency_ms": float(np.percentile(lat, 99)),
            "iterations": sum(r.iterations for r in out), "max_relres": max(r.relres for r in out + out_p),
            "server_phase_s_at_capacity": timing, "host_share_of_wall": host / wall,
            "server_stats": {k: v for k, v in srv.stats.items() if k != "spmv_bytes"},
        }
        del srv
        torch.cuda.empty_cache()
    structures = {}
    service = []
    seq = serving.make_stream(n, seed=seed)
    t = time.perf_counter()
    for i, r in enumerate(seq):
        t1 = time.perf_counter()
        serving.solve_one_fp64(r, structures)
        service.append(time.perf_counter() - t1)
        if (i + 1) % 256 == 0:
            progress(f"served stream one at a time: {i + 1} of {n}")
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t
    progress(f"served stream one at a time: {n / wall_seq:.1f} solves/s")
    service = np.array(service) * 1e3
    # roofline of the grouped product: the first 256 requests served without graph
    # replay (every launch carries HIP events) and without the structure prefetch thread
    # (whose kernels would share the GPU with the timed launches), algorithmic bytes
    # counted per segment iteration by the server
    nr = min(n, 256)
    srv_r = serving.Server(hbm_budget=60e9, max_segments=256, chunk=16, use_graph=False, prefetch=0)
    srv_r.run(serving.make_stream(8, seed=seed + 1, lo=8, hi=16))
    srv_r.stats["spmv_bytes"] = 0.0
    _lib.profile_enable(True)
    _lib.profile_reset()
    srv_r.run(serving.make_stream(nr, seed=seed))
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    n_g, ms_g = prof.get("grouped_spmv", (0, 0.0))
    ach = srv_r.stats["spmv_bytes"] / (ms_g * 1e-3) / 1e9 if ms_g > 0 else 0.0
    del srv_r
    seq_rate = n / wall_seq
    main = pol_out.get("auto") or next(iter(pol_out.values()))
    res = {
        "workload": f"served stream: {n} requests, 70% diffusion / 30% elasticity Q1 boxes N~U{{8..64}}^3, "
                    "lognormal(0,1) coefficients, Jacobi CG to a 1e-8 FP64 residual, grouped CG + HBM-pool "
                    "admission on device-built box structures; policies: auto (the single-solve rule per request: "
                    "FP64 on boxes this size), split (diffusion FP32 iterations with FP64 reliable updates, "
                    "elasticity FP64), mixed (every request FP32 + FP64 reliable updates), fp64",
        "requests": n, "elements_total": sum(int(np.prod(r.n_elem)) for r in seq),
        "capacity_solves_per_s": main["capacity_solves_per_s"],
        "p50_latency_ms": main["p50_latency_ms"], "p99_latency_ms": main["p99_latency_ms"],
        "policies": pol_out,
        "one_at_a_time_fp64_solves_per_s": seq_rate,
        "one_at_a_time_p50_service_ms": float(np.percentile(service, 50)),
        "one_at_a_time_p99_service_ms": float(np.percentile(service, 99)),
        "speedup_vs_one_at_a_time": {p: v["capacity_solves_per_s"] / seq_rate for p, v in pol_out.items()},
        "roofline": {"kernel": "gcg_spmv (grouped SpMV over the admitted requests + p/q update + p.q), first "
                               f"{nr} requests, policy auto", "bound": "hbm",
                     "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "launches": n_g, "avg_launch_ms": ms_g / max(n_g, 1)},
    }
    if "mixed" in pol_out and "fp64" in pol_out:
        res["mixed_vs_fp64_capacity"] = pol_out["mixed"]["capacity_solves_per_s"] / \
            pol_out["fp64"]["capacity_solves_per_s"]
    return res


# ------------------------------------------------------------------ main
def main_leg(args):
    from iron_amd import _lib, fem
    P = build_problem(args)
    red, csr = P["red"], P["csr"]
    E = P["m"].num_elements
    for _ in range(args.warmup):
        step(P, True, args.inner_rtol)
    _lib.profile_enable(True)
    _lib.profile_reset()
    wall, lat, infos = timed_steps(P, args.steps, True, args.inner_rtol)
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    prof32 = fp32_assembly_profile(P)
    # plain FP64 drop-in, same problem, same GPU
    step(P, False, args.inner_rtol)
    _lib.profile_enable(True)
    _lib.profile_reset()
    wall64, lat64, infos64 = timed_steps(P, max(1, min(args.steps, 3)), False, args.inner_rtol)
    prof64 = _lib.profile_read()
    _lib.profile_enable(False)

    ms_step = 1e3 * wall / args.steps
    lat_ms = np.array(lat) * 1e3
    n_cg, ms_cg = prof.get("cg_spmv_f32", (0, 0.0))
    nnz_f, n_f = red.nnz, red.n_free
    # algorithmic bytes of one fused SpMV launch (FP32): values + columns of the
    # solver matrix, row pointers, z read once, p and q read and written
    bytes_spmv32 = 8.0 * nnz_f + 24.0 * n_f + 4.0
    avg_cg_ms = ms_cg / max(n_cg, 1)
    achieved = bytes_spmv32 / (avg_cg_ms * 1e-3) / 1e9 if n_cg else 0.0
    # calibrated HBM bytes per launch (rocprofv3 counters), for this workload and storage only
    traffic = measured_traffic() if (args.workload == "diffusion48" and args.format == "sell") else None
    n_a32, ms_a32 = prof32.get("assemble_scalar_f32", (0, 0.0))
    n_a64, ms_a64 = prof.get("assemble_scalar", (0, 0.0))
    asm32_ms = ms_a32 / max(n_a32, 1)
    asm64_ms = ms_a64 / max(n_a64, 1)
    pick = fem.auto_precision(1, P["S"].stored)
    out = {
        "metric": "elems/s assembled; solves/s mixed vs FP64 at 1e-8 rel resid; p50/p99",
        "value": E * args.steps / wall,
        "unit": "elems/s (assembled, solved to 1e-8 and back-substituted, full pass)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (generated mesh, lognormal conductivity from a fixed seed)",
        "config": {"workload": WORKLOAD_TEXT[args.workload].format(seed=args.seed),
                   "elements": E, "dofs": csr.n, "free_dofs": n_f, "solver_nnz": nnz_f, "eq_nnz": csr.nnz},
        "solves_per_s_mixed": args.steps / wall,
        "solves_per_s_fp64": len(lat64) / wall64,
        "api_precision": pick,
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "p50_ms_fp64": float(np.percentile(np.array(lat64) * 1e3, 50)),
        "assembly_elems_per_s_f32": E / (asm32_ms * 1e-3) if asm32_ms > 0 else None,
        "assembly_elems_per_s_f64": E / (asm64_ms * 1e-3) if asm64_ms > 0 else None,
        "dof_per_s": csr.n * args.steps / wall,
        "iterations_mixed_inner": infos[-1].iterations, "refinement_steps": infos[-1].outer,
        "iterations_fp64": infos64[-1].iterations,
        "relres_mixed": infos[-1].relres, "relres_fp64": infos64[-1].relres,
        "roofline": {"kernel": "cg_spmv<float> (fused SpMV + p/q update + p.q)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_spmv32, "avg_launch_ms": avg_cg_ms, "launches": n_cg},
        "kernel_ms": {k: v for k, v in prof.items() if v[0]},
        "kernel_ms_fp64": {k: v for k, v in prof64.items() if v[0]},
        "solver_format": args.format, "sell_padding": getattr(P["S"], "padding", 0.0),
        "sell_stored": getattr(P["S"], "stored", 0),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_scalar(P, infos64[-1].iterations, args.cpu_seconds)
        nthr = host_threads()
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline_scalar(P, infos64[-1].iterations, 0.5 * args.cpu_seconds,
                                                                nthr)
    del P
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(served requests would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, str(ROOT))
    import torch
    torch.cuda.init()
    if args.calibrate_copy:
        a = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
        bcopy = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=bcopy)
        torch.cuda.synchronize()
        del a, bcopy
    legs = set(args.legs.split(","))
    cpu = 0.0 if args.no_cpu_baseline else args.cpu_seconds
    progress(f"legs {sorted(legs)}")
    out = main_leg(args) if "main" in legs else {"legs": sorted(legs)}
    progress("main leg done")
    torch.cuda.empty_cache()
    if "mr" in legs and not args.no_mr:
        out["mooney_rivlin"] = mr_leg(2, min(cpu, 8.0))
        progress("Mooney-Rivlin leg done")
        torch.cuda.empty_cache()
    if "elastic" in legs:
        out["elasticity96"] = elastic_leg(cpu_seconds=min(cpu, 10.0))
        progress("elasticity leg done")
        torch.cuda.empty_cache()
    if "laplace" in legs:
        out["laplace16"] = laplace_leg(args, cpu_seconds=min(cpu, 4.0))
        progress("Laplace leg done")
    if "slab" in legs:
        out["diffusion48_slab"] = slab_leg(args)
        progress("slab leg done")
    if "serve" in legs and args.serve_requests > 0:
        out["served_stream"] = serve_leg(args.serve_requests, args.serve_policies.split(","))
        progress("served-stream leg done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
