"""ctypes binding of libmi_fe.so (declared in include/mife.h).

The library is required: importing this module on a machine without the built
library raises, and there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libmi_fe.so"

MIFE_OK = 0
MIFE_ERR_NOT_CONVERGED = 4
MIFE_F64 = 0
MIFE_F32 = 1


class CMFEError(RuntimeError):
    """Raised on a nonzero status, as the reference bindings raise CMFEError
    (bindings/python/opencmiss/iron/_utils.in.py:66-77)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (error code {code})")
        self.code = code


class NotConvergedError(CMFEError):
    pass


class MatrixDesc(C.Structure):
    """mife_matrix: a square solver matrix in CSR, SELL-64 or node-block SELL3 storage."""
    _fields_ = [("n", C.c_int), ("format", C.c_int), ("ptr", C.c_void_p), ("col", C.c_void_p),
                ("perm", C.c_void_p), ("n_slices", C.c_int), ("nnz", C.c_int64), ("col_base", C.c_void_p),
                ("slice_order", C.c_void_p), ("order_span", C.c_int)]


MIFE_FMT_CSR = 0
MIFE_FMT_SELL = 1
MIFE_FMT_SELL3 = 3


class Segment(C.Structure):
    """mife_segment: one system of a grouped solve."""
    _fields_ = [("n", C.c_int32), ("maxit", C.c_int32), ("slot", C.c_int32), ("format", C.c_int32),
                ("row_off", C.c_int64), ("rp_off", C.c_int64), ("nz_off", C.c_int64), ("rtol", C.c_double),
                ("n_slices", C.c_int32), ("prec", C.c_int32), ("cb_off", C.c_int64)]


class Pools(C.Structure):
    """mife_pools: shared storage of grouped segments."""
    _fields_ = [("rp", C.c_void_p), ("col", C.c_void_p), ("col16", C.c_void_p), ("cbase", C.c_void_p),
                ("val", C.c_void_p), ("val32", C.c_void_p)]


class ElementDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_nodes", C.c_int), ("n_gauss", C.c_int),
                ("dN", C.c_void_p), ("N", C.c_void_p), ("w", C.c_void_p)]


class BoxDesc(C.Structure):
    """mife_box_desc: a generated box, its order, components and Dirichlet x-planes."""
    _fields_ = [("n_elem", C.c_int32 * 3), ("order", C.c_int32), ("n_comp", C.c_int32), ("fix_lo", C.c_int32),
                ("fix_hi", C.c_int32), ("block", C.c_int32), ("extent", C.c_double * 3), ("u_lo", C.c_double),
                ("u_hi", C.c_double)]


class BoxCounts(C.Structure):
    """mife_box_counts: sizes of a box's structures (closed forms, host)."""
    _fields_ = [("n_nodes", C.c_int64), ("n_elem", C.c_int64), ("nodes_per_elem", C.c_int64), ("n_dofs", C.c_int64),
                ("nnz", C.c_int64), ("n_free", C.c_int64), ("red_nnz", C.c_int64), ("n_slices", C.c_int64),
                ("width", C.c_int64), ("stored", C.c_int64), ("col16", C.c_int32), ("block", C.c_int32)]


class BoxArrays(C.Structure):
    """mife_box_arrays: device arrays mife_box_build writes."""
    _fields_ = [(name, C.c_void_p) for name in ("coords", "elem_nodes", "row_ptr", "col", "free_idx", "new_idx", "u0",
                                                 "slice_ptr", "sell_col", "sell_col16", "col_base", "sell_src")]


MIFE_JOB_SCALAR = 0
MIFE_JOB_ELASTIC = 1


class AssemblyJob(C.Structure):
    """mife_assembly_job: one mesh of a batched assembly."""
    _fields_ = [("kind", C.c_int32), ("n_elem", C.c_int32), ("coords", C.c_void_p), ("elem_nodes", C.c_void_p),
                ("elem_dofs", C.c_void_p), ("coef", C.c_void_p), ("source", C.c_void_p), ("pos", C.c_void_p),
                ("order", C.c_void_p), ("colour_ptr_host", C.c_void_p), ("n_colours", C.c_int32),
                ("pad", C.c_int32), ("vals", C.c_void_p), ("src_vec", C.c_void_p)]


class MrDesc(C.Structure):
    """mife_mr_desc: displacement basis, pressure DOFs per element, formulation."""
    _fields_ = [("u", C.POINTER(ElementDesc)), ("n_pressure", C.c_int), ("Np", C.c_void_p),
                ("formulation", C.c_int)]


class MrPlan(C.Structure):
    """mife_mr_plan: per CSR entry and per DOF, the element stage slots that add into it."""
    _fields_ = [("nnz", C.c_int64), ("n_dofs", C.c_int), ("pad", C.c_int), ("jac_ptr", C.c_void_p),
                ("jac_list", C.c_void_p), ("dof_ptr", C.c_void_p), ("dof_list", C.c_void_p)]


MIFE_MR_DEVIATORIC = 0
MIFE_MR_CAUCHY = 1
MIFE_MR_NEARLY_INCOMPRESSIBLE = 2

_lib = None

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_D = C.c_double
_SZ = C.c_size_t
_SIG = {
    "mife_version": (C.c_char_p, []),
    "mife_last_error": (_I, [C.c_char_p, _I]),
    "mife_init": (_I, [_I]),
    "mife_csr_pattern_work_bytes": (C.c_size_t, [_I, _I, _I]),
    "mife_csr_pattern": (_I, [_P, _I, _I, _I, _P, _P, C.POINTER(_I64), _P, _P]),
    "mife_scatter_map": (_I, [_P, _I, _I, _P, _P, _P, _P]),
    "mife_colour_elements": (_I, [_P, _I, _I, _I, _P, C.POINTER(_I)]),
    "mife_reduce_pattern": (_I, [_I, _P, _P, _P, _I, _P, _P, _P, C.POINTER(_I64), _P, _P]),
    "mife_assemble_scalar": (_I, [C.POINTER(ElementDesc), _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _P]),
    "mife_assemble_elastic": (_I, [C.POINTER(ElementDesc), _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P]),
    "mife_assemble_grid_stage_bytes": (C.c_size_t, [_I, _I, _I, _I]),
    "mife_assemble_grid": (_I, [C.POINTER(ElementDesc), _P, _I, _P, _P, _P, _P, _P, _I64, _I, _P, _P, _P, _SZ,
                                _P]),
    "mife_box_sizes": (_I, [C.POINTER(BoxDesc), C.POINTER(BoxCounts)]),
    "mife_box_build": (_I, [C.POINTER(BoxDesc), C.POINTER(BoxArrays), _P]),
    "mife_spmv": (_I, [C.POINTER(MatrixDesc), _P, _P, _P, _I, _P]),
    "mife_sell_permutation": (_I, [_I, _P, _P, _I, _P, _P]),
    "mife_sell_build": (_I, [_I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
    "mife_sell_compress": (_I, [_I, _P, _P, _P, _P, C.POINTER(_I), _P]),
    "mife_sell_order_len": (_I, [_I]),
    "mife_sell_order": (_I, [_I, _P, _P, _P, C.POINTER(_I), _P]),
    "mife_gather_values": (_I, [_I64, _P, _P, _I, _P, _I, _P]),
    "mife_gather_values2": (_I, [_I64, _P, _P, _P, _P, _P]),
    "mife_solver_rhs": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mife_field_update": (_I, [_I, _P, _P, _P, _P]),
    "mife_couple_pattern_work_bytes": (C.c_size_t, [_I, _I64, _I]),
    "mife_couple_pattern": (_I, [_I, _I64, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_I64),
                                 C.POINTER(_I64), _P, _P]),
    "mife_couple_values": (_I, [_I64, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P]),
    "mife_solver_rhs_coupled": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mife_field_update_coupled": (_I, [_I, _P, _P, _P, _P, _P]),
    "mife_backsubstitute": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _P, _P]),
    "mife_solver_work_bytes": (C.c_size_t, [_I, _I]),
    "mife_cg": (_I, [C.POINTER(MatrixDesc), _P, _P, _P, _I, _D, _I, _P, _SZ, C.POINTER(_I), C.POINTER(_D), _P]),
    "mife_cg_noalloc": (_I, [C.POINTER(MatrixDesc), _P, _P, _P, _I, _D, _I, C.POINTER(_I), C.POINTER(_D), _P]),
    "mife_refine": (_I, [C.POINTER(MatrixDesc), _P, _P, _P, _P, _D, _D, _I, _P, _SZ, C.POINTER(_I), C.POINTER(_I),
                         C.POINTER(_D), _P]),
    "mife_gmres": (_I, [C.POINTER(MatrixDesc), _P, _P, _P, _P, _D, _I, _I, _P, _SZ, C.POINTER(_I), C.POINTER(_D),
                        _P]),
    "mife_minres": (_I, [C.POINTER(MatrixDesc), _P, _I, _P, _P, _P, _D, _I, _P, _SZ, C.POINTER(_I), C.POINTER(_D),
                         _P]),
    "mife_assemble_batch_work_bytes": (C.c_size_t, [_I, _I]),
    "mife_assemble_batch": (_I, [C.POINTER(ElementDesc), C.POINTER(ElementDesc), _I, C.POINTER(AssemblyJob), _I, _I,
                                 _P, _P, _P]),
    "mife_mr_assemble": (_I, [C.POINTER(MrDesc), _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
    "mife_mr_plan_work_bytes": (C.c_size_t, [_I, _I, _I64, _I]),
    "mife_mr_plan_build": (_I, [_I, _I, _P, _I64, _P, _I, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "mife_mr_stage_bytes": (C.c_size_t, [C.POINTER(MrDesc), _I]),
    "mife_mr_assemble_staged": (_I, [C.POINTER(MrDesc), _I, _P, _P, _P, _P, _P, _P, C.POINTER(MrPlan), _P, _P,
                                     _P, _P, _P]),
    "mife_grouped_state_bytes": (C.c_size_t, [_I]),
    "mife_grouped_work_bytes": (C.c_size_t, [_I, _I]),
    "mife_grouped_blocks": (_I, [_I]),
    "mife_grouped_setup": (_I, [_P, _I, _I, _I, C.POINTER(_I), _P, _P, _P]),
    "mife_grouped_cg_init": (_I, [_I, _I, _I, _I, C.POINTER(Pools), _P, _P, _P, _P, _I64, _P, _P, _P]),
    "mife_grouped_cg_iterate": (_I, [_I, _I, _I, _I, C.POINTER(Pools), _P, _P, _P, _P, _I64, _I, _I, _D, _P, _P,
                                     _P]),
    "mife_grouped_cg_residual": (_I, [_I, _I, _I, _I, C.POINTER(Pools), _P, _P, _P, _P, _P]),
    "mife_grouped_cg_status": (_I, [_I, _P, _P, _P, _P, _P]),
    "mife_profile_enable": (_I, [_I]),
    "mife_graphs_enable": (_I, [_I]),
    "mife_gmres_cgs_refinement": (_I, [_I]),
    "mife_graphs_cached": (_I, []),
    "mife_profile_reset": (_I, []),
    "mife_profile_read": (_I, [_I, C.c_char_p, _P, _P, C.POINTER(_I)]),
}


def lib():
    """Load (once) and return the library handle."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m iron_amd.build` (hipcc, gfx950)")
        # torch brings its own HIP runtime with the same soname (libamdhip64.so.7): load
        # it first, so that the library binds to that one runtime.  Loaded the other way
        # round the process holds two runtimes, and device memory from one is foreign to
        # kernels launched through the other.
        import torch  # noqa: F401
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def exported_symbols():
    return list(_SIG)


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().mife_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code: int):
    if code == MIFE_OK:
        return
    msg = last_error()
    if code == MIFE_ERR_NOT_CONVERGED:
        raise NotConvergedError(code, msg)
    raise CMFEError(code, msg)


def ptr(t) -> int:
    """Device (or host) address of a torch tensor / numpy array, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def call(name: str, *args):
    check(getattr(lib(), name)(*args))


def graphs_enable(on: bool = True):
    """HIP-graph replay of CG iteration chunks on small systems (default on)."""
    call("mife_graphs_enable", 1 if on else 0)


def gmres_cgs_refinement(mode: int):
    """GMRES classical Gram-Schmidt refinement: 0 never (one pass), 1 always."""
    call("mife_gmres_cgs_refinement", int(mode))


def profile_enable(on: bool = True):
    call("mife_profile_enable", 1 if on else 0)


def profile_reset():
    call("mife_profile_reset")


def profile_read() -> dict:
    """{kernel: (launches, total_ms)} of the HIP-event timings since the last reset."""
    import numpy as np
    n_max = 32
    names = C.create_string_buffer(48 * n_max)
    counts = np.zeros(n_max, dtype=np.int32)
    ms = np.zeros(n_max, dtype=np.float64)
    n = C.c_int(0)
    call("mife_profile_read", n_max, names, counts.ctypes.data, ms.ctypes.data, C.byref(n))
    out = {}
    for k in range(n.value):
        name = names.raw[48 * k: 48 * (k + 1)].split(b"\0", 1)[0].decode()
        out[name] = (int(counts[k]), float(ms[k]))
    return out
