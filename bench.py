#!/usr/bin/env python
"""Benchmark of iron's element-assembly + sparse-solve hot path on one MI355X.

Workload (default ``diffusion48``): 3-D Poisson-type diffusion on a 48x48x48
triquadratic Lagrange hex mesh (110,592 elements, 97^3 nodes, 3x3x3 Gauss),
per-element conductivity lognormal(0,1) from a fixed seed, source f = 1, u = 0 on
every boundary face.  One step is one full pass of the path:
  FP32 (served inner operator) and FP64 (outer operator) element integration with
  coloured scatter into the CSR -> K_ff gather + RHS -> FP32 Jacobi CG inside FP64
  iterative refinement to ||b - A x|| <= 1e-8 ||b|| -> field update.
The plain FP64 drop-in (FP64 assembly + FP64 Jacobi CG to 1e-8) is timed in the
same run, and the C restatement of iron's CPU algorithm on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload diffusion48|laplace16] [--legs main,mr,elastic,laplace,slab,serve]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="diffusion48", choices=["diffusion48", "laplace16"])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--format", default="sell", choices=["sell", "csr"], help="solver-matrix storage")
    ap.add_argument("--inner-rtol", type=float, default=0.1,
                    help="FP32 residual reduction between FP64 residual replacements")
    ap.add_argument("--no-mr", action="store_true", help="skip the Mooney-Rivlin (12^3 Q2/Q1) leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run three streaming passes y = 1.0 * x over 256 MiB FP32 arrays (reads and "
                         "writes 268,435,456 bytes each: the reference kernel for HBM byte counters)")
    ap.add_argument("--serve-requests", type=int, default=128,
                    help="requests of the served-stream leg (the full stream has 1024); 0 skips it")
    ap.add_argument("--legs", default="main,mr,elastic,laplace,slab,serve",
                    help="comma-separated legs to run: main (the workload above), mr, elastic (96x32x32 "
                         "elasticity), laplace (16^3 Laplace), slab (one eighth of the 48^3 case), serve")
    return ap.parse_args()


TRAFFIC_FILE = ROOT / "profiles" / "cg_spmv_diffusion48_q2.md"
WORKLOAD_TEXT = {
    "diffusion48": "diffusion48: 48^3 Q2 hex, 3x3x3 Gauss, kappa~lognormal(0,1) seed {seed}, f=1, u=0 on all "
                   "faces, FP32 inner CG + FP64 refinement to 1e-8",
    "diffusion48slab": "diffusion48 slab: 6x48x48 Q2 hex on [0,1/8]x[0,1]^2 (one of eight x-slabs of the 48^3 "
                       "case), 3x3x3 Gauss, kappa~lognormal(0,1) seed {seed}, f=1, u=0 on all faces, FP32 inner CG "
                       "+ FP64 refinement to 1e-8",
    "laplace16": "laplace16: 16^3 Q1 hex, 2x2x2 Gauss, kappa=1, u=0 at x=0 and u=1 at x=1, FP32 inner CG + "
                 "FP64 refinement to 1e-10",
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
    asm = fem.ScalarAssembler(b, m.coords, m.elem_nodes[0], csr)
    red = fem.DirichletReduction(csr, fixed)
    S = fem.SolverMatrix(red, args.format)
    P = dict(m=m, dm=dm, basis=b, csr=csr, asm=asm, red=red, S=S, rtol=rtol, fixed=fixed, ufix=ufix,
             kappa=kappa.to(dev), source=None if source is None else source.to(dev), kappa_h=kappa.numpy(),
             ws=fem.KrylovWorkspace(red.n_free, dev))
    P["v64"] = csr.zeros(fem.F64)
    P["v32"] = csr.zeros(fem.F32)
    P["src"] = torch.zeros(dm.num_dofs, dtype=torch.float64, device=dev)
    P["u"] = torch.as_tensor(np.where(fixed, ufix, 0.0), device=dev)
    P["sv64"] = torch.empty(max(S.stored, 1), dtype=torch.float64, device=dev)
    P["sv32"] = torch.empty(max(S.stored, 1), dtype=torch.float32, device=dev)
    P["b"] = torch.empty(red.n_free, dtype=torch.float64, device=dev)
    P["x"] = torch.zeros(red.n_free, dtype=torch.float64, device=dev)
    return P


def step(P, mixed: bool, inner_rtol: float):
    from iron_amd import fem
    v64, v32, src = P["v64"], P["v32"], P["src"]
    v64.zero_()
    src.zero_()
    P["asm"].assemble(v64, kappa=P["kappa"], source=P["source"], src_vec=src if P["source"] is not None else None,
                      prec=fem.F64)
    red, S = P["red"], P["S"]
    S.gather(v64, fem.F64, out=P["sv64"])
    red.rhs(v64, P["u"], src=src if P["source"] is not None else None, out=P["b"])
    x = P["x"]
    x.zero_()
    if mixed:
        v32.zero_()
        P["asm"].assemble(v32, kappa=P["kappa"], prec=fem.F32)
        S.gather(v32, fem.F32, out=P["sv32"])
        info = fem.refine(S, P["sv64"], P["sv32"], P["b"], x, P["rtol"], 200000, P["ws"], inner_rtol=inner_rtol)
    else:
        info = fem.cg(S, P["sv64"], P["b"], x, P["rtol"], 200000, P["ws"], prec=fem.F64)
    red.field_update(x, P["u"])
    return info


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


def host_threads():
    """Host threads for the all-cores CPU baseline: OMP_NUM_THREADS when set (the
    host-core share given to this GPU), else the CPUs this process may run on."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return len(os.sched_getaffinity(0))


def cpu_baseline(P, fp64_iters, budget_s, threads=1):
    """Times the C restatement of iron's algorithm (oracle/liboracle_fe.so) over a
    bounded sample and scales it to one full pass of the workload: on one core (iron's
    per-rank serial loop), or on `threads` OpenMP threads (elements shared out with
    atomic adds, CG rows and dot products split) as a proxy for iron on that many
    ranks."""
    import ctypes as C
    sys.path.insert(0, str(ROOT))
    so = ROOT / "oracle" / "liboracle_fe.so"
    if not so.exists():
        return None
    lib = C.CDLL(str(so))
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
    P_ = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    lib.oracle_assemble_scalar3.restype = C.c_int
    lib.oracle_assemble_scalar3_par.restype = C.c_int
    nthr = host_threads()

    def assemble(e0, e1, nt):
        args = (E, b.num_nodes, b.num_gauss, P_(dN), P_(N), P_(w), P_(coords), P_(en), P_(ed), P_(kap), P_(f),
                P_(rp), P_(col), P_(vals), P_(src), e0, e1)
        return lib.oracle_assemble_scalar3(*args) if nt == 1 else lib.oracle_assemble_scalar3_par(*args, nt)

    # assembly sample: elements [0, Es) chosen to fit ~40% of the budget
    t = time.perf_counter()
    probe = min(E, 512 * threads)
    assemble(0, probe, threads)
    per_elem = (time.perf_counter() - t) / probe
    Es = int(min(E, max(probe, 0.4 * budget_s / max(per_elem, 1e-9))))
    vals[:] = 0.0
    src[:] = 0.0
    t = time.perf_counter()
    assemble(0, Es, threads)
    t_asm = (time.perf_counter() - t) * (E / Es)
    if Es < E:  # complete the matrix (untimed, all host threads) so the solve runs on the real system
        assemble(Es, E, nthr)
    new_idx = red.new_idx.cpu().numpy().astype(np.int32)
    u = P["u"].cpu().numpy()
    srp = np.zeros(red.n_free + 1, dtype=np.int32)
    scol = np.zeros(max(red.nnz, 1), dtype=np.int32)
    sv = np.zeros(max(red.nnz, 1))
    bb = np.zeros(red.n_free)
    lib.oracle_reduce.restype = C.c_int64
    t = time.perf_counter()
    lib.oracle_reduce(csr.n, P_(rp), P_(col), P_(vals), P_(new_idx), P_(u), P_(src if f is not None else None),
                      P_(srp), P_(scol), P_(sv), P_(bb))
    t_red = time.perf_counter() - t
    x = np.zeros(red.n_free)
    rr = C.c_double(0.0)
    lib.oracle_cg.restype = C.c_int
    lib.oracle_cg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                              C.c_int, C.c_int, C.POINTER(C.c_double)]
    t = time.perf_counter()
    it_probe = 5
    lib.oracle_cg(red.n_free, P_(srp), P_(scol), P_(sv), P_(bb), P_(x), 0.0, it_probe, threads, C.byref(rr))
    per_it = (time.perf_counter() - t) / it_probe
    n_it = int(max(it_probe, min(fp64_iters, 0.5 * budget_s / max(per_it, 1e-9))))
    x[:] = 0.0
    t = time.perf_counter()
    lib.oracle_cg(red.n_free, P_(srp), P_(scol), P_(sv), P_(bb), P_(x), 0.0, n_it, threads, C.byref(rr))
    per_it = (time.perf_counter() - t) / n_it
    t_solve = per_it * fp64_iters
    t_total = t_asm + t_red + t_solve
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": E / t_total, "unit": "elems/s", "cores": threads, "kind": "port",
        "sample": (f"C restatement of iron's FP64 path on {threads} thread(s) ({cpu}): element loop with bisection "
                   f"CSR add"
                   f" over {Es} of {E} elements (x{E / Es:.2f}), K_ff/RHS reduction in full (serial), {n_it} Jacobi-CG "
                   f"iterations scaled to the {fp64_iters} iterations the FP64 solve needs"),
        "assembly_s": t_asm, "reduce_s": t_red, "cg_iter_ms": per_it * 1e3, "solve_s": t_solve,
        "solves_per_s": 1.0 / t_total,
    }


def mr_flops_per_element(nn, npd, ng):
    """Algorithmic FP64 flops of one Mooney-Rivlin element (residual + exact Jacobian),
    counted from the kernel's loops: per Gauss point 72 nn^2 (displacement block:
    (3nn)^2 entries x 3 FMAs + weight), 54 nn npd (pressure coupling), 216 nn (metrics,
    reference gradients, T = GX . A) and about 4,200 for kinematics, stress and the
    81-entry tangent."""
    return ng * (72 * nn * nn + 54 * nn * npd + 216 * nn + 4200)


def mr_leg(runs, cpu_seconds):
    """12^3 triquadratic/trilinear Mooney-Rivlin cube (c1=2, c2=6), rollers on x=0, y=0,
    z=0, x=1 face stretched to 1.2 in 10 load increments, Newton to a relative residual
    of 1e-8 per increment (preconditioned MINRES steps, Eisenstat-Walker tolerances)."""
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
    opts = fe.NewtonOptions(atol=1e-12, rtol=1e-8, stol=0.0, linear_rtol="ew")

    def run():
        x = torch.as_tensor(x0, device=sysg.device)
        infos = fe.load_increments(newton, x, 10, fixed, target, opts=opts)
        torch.cuda.synchronize()
        return x, infos

    run()
    _lib.profile_enable(True)
    _lib.profile_reset()
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        x, infos = run()
        times.append(time.perf_counter() - t)
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    xe = X * np.array([1.2, 1.2 ** -0.5, 1.2 ** -0.5])
    err = float(np.abs(x.cpu().numpy()[:3 * nu].reshape(3, -1).T - xe).max())
    n_mr, ms_mr = prof.get("finite_elasticity", (0, 0.0))
    avg_mr = ms_mr / max(n_mr, 1)
    E = rm.num_elements
    fl = mr_flops_per_element(27, 8, 27) * E
    out = {
        "workload": "mr12: 12^3 Q2/Q1 Mooney-Rivlin (deviatoric, c1=2, c2=6), rollers + x-stretch to 1.2 in 10 "
                    "increments, Newton to 1e-8 per increment, MINRES steps with Eisenstat-Walker tolerances",
        "elements": E, "dofs": sysg.ndofs, "free_dofs": newton.red.n_free, "jacobian_nnz": sysg.csr.nnz,
        "solve_s": float(np.median(times)), "solves_per_s": 1.0 / float(np.median(times)),
        "newton_iterations": sum(i.iterations for i in infos),
        "minres_iterations": sum(i.linear_iterations for i in infos),
        "max_node_error_vs_closed_form": err,
        "assembly_ms": avg_mr, "assemblies": n_mr,
        "assembly_elems_per_s": E / (avg_mr * 1e-3) if avg_mr > 0 else None,
        "roofline": {"kernel": "mr_element_kernel<27,8,27> (residual + exact Jacobian, FP64)", "bound": "fp64",
                     "achieved": fl / (avg_mr * 1e-3) / 1e12 if avg_mr > 0 else None, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": (fl / (avg_mr * 1e-3) / 1e12) / 78.6 if avg_mr > 0 else None,
                     "flops_per_element": mr_flops_per_element(27, 8, 27)},
        "kernel_ms": {k: v for k, v in prof.items() if v[0]},
    }
    # the same load path with GMRES steps (restart 100, pressure-block preconditioner,
    # FP64 values), each linear solve to 1e-8
    opts_g = fe.NewtonOptions(atol=1e-12, rtol=1e-8, stol=0.0, linear_rtol=1e-8, linear="gmres")
    x = torch.as_tensor(x0, device=sysg.device)
    t = time.perf_counter()
    infos_g = fe.load_increments(newton, x, 10, fixed, target, opts=opts_g)
    torch.cuda.synchronize()
    out["gmres_1e8"] = {"solve_s": time.perf_counter() - t, "newton_iterations": sum(i.iterations for i in infos_g),
                        "gmres_iterations": sum(i.linear_iterations for i in infos_g),
                        "max_node_error_vs_closed_form":
                            float(np.abs(x.cpu().numpy()[:3 * nu].reshape(3, -1).T - xe).max())}
    if cpu_seconds > 0:
        sys.path.insert(0, str(ROOT))
        from oracle import mr_oracle as mo
        from oracle.fe_oracle import Basis, regular_mesh
        om = regular_mesh((12, 12, 12), [Basis(3, 2, 3), Basis(3, 1, 3)], (1.0, 1.0, 1.0))
        prob = mo.MRProblem(om, False, np.full(E, 2.0), np.full(E, 6.0))
        xs = prob.initial_state(0.0)
        geo_e, dep_e, pe = prob.element_arrays(xs)
        ns = 16
        t = time.perf_counter()
        k = 0
        while time.perf_counter() - t < 0.5 * cpu_seconds and k + ns <= E:
            mo.mr_element_residual(geo_e[k:k + ns], dep_e[k:k + ns], pe[k:k + ns], prob.bu, prob.bp,
                                   prob.c1[k:k + ns], prob.c2[k:k + ns])
            mo.mr_element_jacobian(geo_e[k:k + ns], dep_e[k:k + ns], pe[k:k + ns], prob.bu, prob.bp,
                                   prob.c1[k:k + ns], prob.c2[k:k + ns])
            k += ns
        per = (time.perf_counter() - t) / max(k, 1)
        out["cpu_baseline"] = {"value": 1.0 / per, "unit": "elems/s (residual + Jacobian)", "cores": 1,
                               "kind": "port",
                               "sample": f"NumPy FP64 oracle element residual + exact Jacobian on {k} of {E} "
                                         f"elements (vectorised over {ns} elements per call)"}
    return out


def elastic_leg(runs=3):
    """96x32x32 trilinear linear elasticity on [0,3]x[0,1]^2: lognormal Young's modulus
    10 exp(0.3 N(0,1)) per element (seed 7), Poisson ratio 0.3, clamp at x = 0,
    consistent traction t_z = -1 on x = 3.  One pass: FP64 and FP32 (MFMA) element
    stiffness, solver values and RHS, FP32 Jacobi CG with FP64 refinement to a 1e-8
    relative residual (rows node by node, 16-bit SELL columns), field update.  Also
    the FP64 drop-in CG on the same system."""
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
    mat = torch.as_tensor(np.ascontiguousarray(np.stack([Ee, Ee, Ee, np.full(E, 0.3), np.full(E, 0.3),
                                                         np.full(E, 0.3)], axis=1)), device=dev)
    csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, dev, colours=mesh.structured_colours(m))
    asm = fem.ElasticAssembler(b, X, m.elem_nodes[0], csr)
    clamp = X[:, 0] == 0.0
    red = fem.DirichletReduction(csr, np.concatenate([clamp, clamp, clamp]))
    S = fem.SolverMatrix(red, interleave=(3, N))
    tip = np.nonzero(X[:, 0] == 3.0)[0]
    ft = torch.as_tensor(engine.neumann_integrate(m, dm, [b] * 3, {int(2 * N + i): -1.0 for i in tip}), device=dev)
    v, v32 = csr.zeros(fem.F64), csr.zeros(fem.F32)
    sv64 = torch.empty(max(S.stored, 1), dtype=torch.float64, device=dev)
    sv32 = torch.empty(max(S.stored, 1), dtype=torch.float32, device=dev)
    u = torch.zeros(csr.n, dtype=torch.float64, device=dev)
    bvec = torch.empty(red.n_free, dtype=torch.float64, device=dev)
    x = torch.zeros(red.n_free, dtype=torch.float64, device=dev)
    ws = fem.KrylovWorkspace(red.n_free, dev)

    def one(mixed=True):
        v.zero_()
        asm.assemble(v, mat)
        S.gather(v, fem.F64, out=sv64)
        u.zero_()
        red.rhs(v, u, rhs=ft, out=bvec)
        x.zero_()
        if mixed:
            v32.zero_()
            asm.assemble(v32, mat, prec=fem.F32)
            S.gather(v32, fem.F32, out=sv32)
            info = fem.refine(S, sv64, sv32, bvec, x, 1e-8, 200000, ws)
        else:
            info = fem.cg(S, sv64, bvec, x, 1e-8, 200000, ws, prec=fem.F64)
        red.field_update(x, u)
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
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    n_s, ms_s = prof.get("cg_spmv_f32", (0, 0.0))
    per = ms_s / max(n_s, 1)
    alg = 8.0 * red.nnz + 24.0 * red.n_free   # values + int32 columns, z, p/q read, p/q written
    t_mx, info_mx = res[True]
    t_64, info_64 = res[False]
    a64 = prof.get("assemble_elastic", (0, 0.0))[1]
    a32 = prof.get("assemble_elastic_f32", (0, 0.0))[1]
    return {
        "workload": "elasticity96: 96x32x32 Q1 hex on [0,3]x[0,1]^2, 2x2x2 Gauss, E = 10 exp(0.3 N(0,1)) per element "
                    "(seed 7), nu = 0.3, clamp at x=0, traction t_z=-1 on x=3; FP32 element stiffness (MFMA) + FP32 "
                    "Jacobi CG with FP64 refinement to 1e-8",
        "elements": E, "free_dofs": red.n_free, "solver_nnz": red.nnz, "sell_padding": S.padding,
        "ms_per_pass": t_mx * 1e3, "elems_per_s": E / t_mx, "iterations_inner": info_mx.iterations,
        "refinement_steps": info_mx.outer, "relres": info_mx.relres,
        "fp64_ms_per_pass": t_64 * 1e3, "fp64_iterations": info_64.iterations, "fp64_relres": info_64.relres,
        "assembly_ms_f64": a64, "assembly_ms_f32": a32,
        "assembly_elems_per_s_f32": E / (a32 * 1e-3) if a32 > 0 else None,
        "roofline": {"kernel": "cg_spmv<float> (fused SpMV + p/q update + p.q, 16-bit columns)", "bound": "hbm",
                     "achieved": alg / (per * 1e-3) / 1e9 if per > 0 else 0.0, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (alg / (per * 1e-3) / 1e9 / HBM_PEAK_GBS) if per > 0 else 0.0,
                     "bytes_per_launch": alg, "avg_launch_ms": per, "launches": n_s},
    }


def laplace_leg(args, runs=5):
    """16^3 trilinear Laplace, u = 0 at x = 0 and u = 1 at x = 1: the full pass (FP64
    assembly, solver values, FP32 CG with FP64 refinement to 1e-10, field update), and
    the FP64 drop-in CG."""
    import argparse
    import torch
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
    torch.cuda.synchronize()
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
    red = P["red"]
    n_s, ms_s = prof.get("cg_spmv_f32", (0, 0.0))
    per = ms_s / max(n_s, 1)
    alg = 8.0 * red.nnz + 24.0 * red.n_free
    E = P["m"].num_elements
    a64 = prof.get("assemble_scalar", (0, 0.0))
    a32 = prof.get("assemble_scalar_f32", (0, 0.0))
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


def serve_leg(n, seed=2024):
    """Served stream: n requests, 70% diffusion / 30% elasticity trilinear boxes with
    N_x, N_y, N_z ~ U{8..64}, lognormal(0,1) coefficients (per-request seeds), BCs of
    the 16^3 Laplace and 96x32x32 elasticity cases, solved to a 1e-8 FP64 residual.
    Measured: the grouped server's capacity (all requests at t=0; FP32 iterations
    with FP64 reliable updates, and FP64 iterations for comparison), then latency at
    Poisson arrivals of 0.7x that capacity (arrival -> solution on the host), and the
    one-at-a-time FP64 path on the same requests."""
    import torch
    from iron_amd import serving
    srv = serving.Server(hbm_budget=60e9, max_segments=256, chunk=16, use_graph=True)
    warm = serving.make_stream(8, seed=seed + 1, lo=8, hi=16)
    srv.run(warm)
    reqs = serving.make_stream(n, seed=seed)
    t = time.perf_counter()
    out = srv.run(reqs)
    torch.cuda.synchronize()
    wall_sat = time.perf_counter() - t
    cap = n / wall_sat
    timing_sat = dict(srv.timing)
    for k in srv.timing:
        srv.timing[k] = 0.0
    reqs_p = serving.make_stream(n, seed=seed, rate=0.7 * cap)
    t = time.perf_counter()
    out_p = srv.run(reqs_p)
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - t
    lat = np.array([r.latency for r in out_p]) * 1e3
    structures = {}
    seq = serving.make_stream(n, seed=seed)
    service = []
    t = time.perf_counter()
    for r in seq:
        t1 = time.perf_counter()
        serving.solve_one_fp64(r, structures)
        service.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t
    service = np.array(service) * 1e3
    elems = sum(int(np.prod(r.n_elem)) for r in reqs)
    stats_sat = dict(srv.stats)
    del srv   # its pools, before the next server takes its own
    # the same stream through FP64 grouped iterations
    srv64 = serving.Server(hbm_budget=60e9, max_segments=256, chunk=16, use_graph=True, precision="fp64")
    srv64.run(serving.make_stream(8, seed=seed + 1, lo=8, hi=16))
    t = time.perf_counter()
    out64 = srv64.run(serving.make_stream(n, seed=seed))
    torch.cuda.synchronize()
    cap64 = n / (time.perf_counter() - t)
    it64 = sum(r.iterations for r in out64)
    del srv64
    # roofline of the grouped product: the same stream served once more without graph
    # replay (so that every launch carries HIP events) and without the structure
    # prefetch thread (whose kernels would share the GPU with the timed launches),
    # algorithmic bytes counted per segment iteration by the server
    from iron_amd import _lib
    srv_r = serving.Server(hbm_budget=60e9, max_segments=256, chunk=16, use_graph=False, prefetch=0)
    srv_r.run(serving.make_stream(8, seed=seed + 1, lo=8, hi=16))
    srv_r.stats["spmv_bytes"] = 0.0
    _lib.profile_enable(True)
    _lib.profile_reset()
    srv_r.run(serving.make_stream(n, seed=seed))
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    n_g, ms_g = prof.get("grouped_spmv", (0, 0.0))
    ach = srv_r.stats["spmv_bytes"] / (ms_g * 1e-3) / 1e9 if ms_g > 0 else 0.0
    roof = {"kernel": "gcg_spmv (grouped SpMV over the admitted requests + p/q update + p.q; FP32 for diffusion "
                      "requests, FP64 for elasticity)",
            "bound": "hbm",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "bytes_total": srv_r.stats["spmv_bytes"], "launches": n_g, "avg_launch_ms": ms_g / max(n_g, 1)}
    return {
        "workload": f"served stream: {n} requests, 70% diffusion / 30% elasticity Q1 boxes N~U{{8..64}}^3, "
                    "lognormal(0,1) coefficients, Jacobi CG to a 1e-8 FP64 residual (diffusion: FP32 iterations with "
                    "FP64 reliable updates; elasticity: FP64), grouped CG + HBM-pool admission",
        "requests": n, "elements_total": elems,
        "capacity_solves_per_s": cap, "elems_per_s_at_capacity": elems / wall_sat,
        "poisson_rate_per_s": 0.7 * cap, "poisson_solves_per_s": n / wall_p,
        "p50_latency_ms": float(np.percentile(lat, 50)), "p99_latency_ms": float(np.percentile(lat, 99)),
        "one_at_a_time_fp64_solves_per_s": n / wall_seq,
        "one_at_a_time_p50_service_ms": float(np.percentile(service, 50)),
        "one_at_a_time_p99_service_ms": float(np.percentile(service, 99)),
        "speedup_vs_one_at_a_time": (n / wall_sat) / (n / wall_seq),
        "grouped_all_fp64_capacity_solves_per_s": cap64,
        "iterations_served_vs_all_fp64": [sum(r.iterations for r in out), it64],
        "replacement_ms": prof.get("grouped_replace", (0, 0.0))[1],
        "max_relres": max(r.relres for r in out + out_p),
        "server_phase_s_at_capacity": timing_sat, "server_stats": stats_sat, "roofline": roof,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(served requests would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, str(ROOT))
    import torch
    from iron_amd import _lib, fem
    torch.cuda.init()
    if args.calibrate_copy:
        a = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
        bcopy = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=bcopy)
        torch.cuda.synchronize()
        del a, bcopy
    legs = set(args.legs.split(","))
    out = main_leg(args) if "main" in legs else {"legs": sorted(legs)}
    torch.cuda.empty_cache()
    if "mr" in legs and not args.no_mr:
        out["mooney_rivlin"] = mr_leg(2, 0.0 if args.no_cpu_baseline else min(args.cpu_seconds, 10.0))
        torch.cuda.empty_cache()
    if "elastic" in legs:
        out["elasticity96"] = elastic_leg()
    if "laplace" in legs:
        out["laplace16"] = laplace_leg(args)
    if "slab" in legs:
        out["diffusion48_slab"] = slab_leg(args)
    if "serve" in legs and args.serve_requests > 0:
        out["served_stream"] = serve_leg(args.serve_requests)
    print(json.dumps(out))


def main_leg(args):
    import torch
    from iron_amd import _lib
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
    n_a32, ms_a32 = prof.get("assemble_scalar_f32", (0, 0.0))
    n_a64, ms_a64 = prof.get("assemble_scalar", (0, 0.0))
    asm32_ms = ms_a32 / max(n_a32, 1)
    asm64_ms = ms_a64 / max(n_a64, 1)
    out = {
        "metric": "elems/s assembled; solves/s mixed vs FP64 at 1e-8 rel resid; p50/p99",
        "value": E * args.steps / wall,
        "unit": "elems/s (assembled and solved to 1e-8, full pass)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (generated mesh, lognormal conductivity from a fixed seed)",
        "config": {"workload": WORKLOAD_TEXT[args.workload].format(seed=args.seed),
                   "elements": E, "dofs": csr.n, "free_dofs": n_f, "solver_nnz": nnz_f, "eq_nnz": csr.nnz},
        "solves_per_s_mixed": args.steps / wall,
        "solves_per_s_fp64": len(lat64) / wall64,
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
        out["cpu_baseline"] = cpu_baseline(P, infos64[-1].iterations, args.cpu_seconds)
        nthr = host_threads()
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline(P, infos64[-1].iterations, 0.5 * args.cpu_seconds, nthr)
    del P
    return out


if __name__ == "__main__":
    main()
