"""The C ABI library loads and exports every entry point include/mife.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "mife.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mife_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mife_assemble_scalar" in names and "mife_cg" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from iron_amd import _lib
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_error_text_roundtrip_without_gpu():
    from iron_amd import _lib
    assert _lib.lib().mife_version().decode().startswith("mi_fe")
    assert _lib.lib().mife_last_error(None, 0) == -1


def test_argument_errors_without_gpu():
    """Argument checks come before any device work and follow the C bindings'
    convention: -1 for a null pointer, positive codes for invalid arguments, with
    the reason in mife_last_error (bindings/generate_bindings/c.py:31-40)."""
    from iron_amd import _lib
    L = _lib.lib()
    M = _lib.MatrixDesc
    # null pointers
    assert L.mife_spmv(None, None, None, None, 0, None) == -1
    assert "null" in _lib.last_error()
    assert L.mife_cg(None, None, None, None, 0, 1e-8, 10, None, 0, None, None, None) == -1
    assert L.mife_sell_order(4, None, None, None, None, None) == -1
    assert L.mife_grouped_setup(None, 0, 1, 1, None, None, None, None) == -1
    pools = _lib.Pools(1, 1, 1, 1, 1, None)   # the FP32 value pool is required
    it = L.mife_grouped_cg_iterate
    assert it(1, 1, 1, 1, ctypes.byref(pools), 1, 1, 1, 1, 4, 4, 4, 0.1, 1, 1, None) == -1
    pools = _lib.Pools(1, 1, 1, 1, 1, 1)
    assert it(1, 1, 1, 1, ctypes.byref(pools), 1, 1, 1, 1, 4, 4, 0, 0.1, 1, 1, None) == 2
    assert "check_every" in _lib.last_error()
    assert it(1, 1, 1, 1, ctypes.byref(pools), 1, 1, 1, 1, 4, 4, 4, 1.5, 1, 1, None) == 2
    seg = _lib.Segment(10, 5, 0, _lib.MIFE_FMT_CSR, 0, 0, 0, 1e-8, 0, 7, -1)   # unknown precision 7
    nb = ctypes.c_int(0)
    stage = bytearray(int(L.mife_grouped_work_bytes(1, 4)))
    buf = (ctypes.c_char * len(stage)).from_buffer(stage)
    assert L.mife_grouped_setup(ctypes.byref(seg), 1, 1, 4, ctypes.byref(nb), buf, buf, None) == 2
    assert "precision" in _lib.last_error()
    # batched assembly: nothing to do, then an unknown job kind and a colour table that
    # does not cover the elements (checked before any device work)
    assert L.mife_assemble_batch(None, None, 0, None, 0, 0, None, None, None) == 0
    job = (_lib.AssemblyJob * 1)()
    job[0].kind, job[0].n_elem = 5, 0
    host = bytearray(int(L.mife_assemble_batch_work_bytes(1, 8)))
    hb = (ctypes.c_char * len(host)).from_buffer(host)
    assert L.mife_assemble_batch(None, None, 1, job, 8, 0, hb, hb, None) == 2
    assert "kind" in _lib.last_error()
    cptr = (ctypes.c_int32 * 3)(0, 4, 9)
    job[0] = _lib.AssemblyJob(0, 10, 1, 1, 1, None, None, 1, 1, ctypes.addressof(cptr), 2, 0, 1, None)
    assert L.mife_assemble_batch(None, None, 1, job, 8, 0, hb, hb, None) == 2
    assert "colour pointers" in _lib.last_error()
    # invalid arguments on a well-formed (host-side) descriptor
    d = M(10, 7, 1, 1, None, 0, 10, None, None, 0)
    assert L.mife_spmv(ctypes.byref(d), 1, 1, 1, 0, None) == 2     # unknown format 7
    assert "format" in _lib.last_error()
    d = M(10, 0, 1, 1, None, 0, 10, None, None, 0)
    assert L.mife_spmv(ctypes.byref(d), 1, 1, 1, 5, None) == 2     # unknown precision 5
    assert "precision" in _lib.last_error()
    assert L.mife_cg(ctypes.byref(d), 1, 1, 1, 0, -1.0, 10, 1, 1 << 30, None, None, None) == 2   # negative rtol
    # a workspace smaller than the solver's layout is refused before any launch,
    # for every solver and every GMRES restart up to 256 (n large enough that the
    # GMRES basis, not the reduction state, sets the size)
    n = 20000
    d = M(n, 0, 1, 1, None, 0, 10, None, None, 0)
    w = L.mife_solver_work_bytes(n, 30)
    assert L.mife_cg(ctypes.byref(d), 1, 1, 1, 0, 1e-8, 10, 1, 1024, None, None, None) == 2
    assert "workspace" in _lib.last_error()
    assert L.mife_refine(ctypes.byref(d), 1, 1, 1, 1, 1e-8, 0.1, 10, 1, 1024, None, None, None, None) == 2
    assert "workspace" in _lib.last_error()
    assert L.mife_minres(ctypes.byref(d), 1, 0, 1, 1, None, 1e-8, 10, 1, 1024, None, None, None) == 2
    assert "workspace" in _lib.last_error()
    for R in (31, 38, 39, 64, 100, 255, 256):
        assert L.mife_gmres(ctypes.byref(d), 1, 1, 1, None, 1e-8, R, 10, 1, w, None, None, None) == 2
        assert "GMRES needs a workspace" in _lib.last_error()
    for R in (30, 100, 256):
        wr = L.mife_solver_work_bytes(n, R)
        assert wr > L.mife_solver_work_bytes(n, R - 1)    # the GMRES basis sets the size here
        # one byte short of the layout at an aligned base is refused (the size is the
        # layout's plus 255 bytes of alignment slack)
        assert L.mife_gmres(ctypes.byref(d), 1, 1, 1, None, 1e-8, R, 10, 256, wr - 256, None, None, None) == 2
        assert "GMRES needs a workspace" in _lib.last_error()
    assert L.mife_gmres(ctypes.byref(d), 1, 1, 1, None, 1e-8, 257, 10, 1, 1 << 40, None, None, None) == 2
    # structured assembly: row count and stage size are checked on the host
    desc = _lib.ElementDesc(3, 27, 27, 1, 1, 1)
    ne = (ctypes.c_int32 * 3)(4, 3, 2)
    rows = 9 * 7 * 5
    need = L.mife_assemble_grid_stage_bytes(24, 27, 1, 0)
    assert need >= 24 * 27 * 27 * 8
    assert L.mife_assemble_grid(ctypes.byref(desc), ne, 1, 1, 1, None, None, 1, rows + 1, 0, 1, None, 1, need,
                                None) == 2
    assert "rows" in _lib.last_error()
    assert L.mife_assemble_grid(ctypes.byref(desc), ne, 1, 1, 1, None, None, 1, rows, 0, 1, None, 1, need - 1,
                                None) == 2
    assert "stage" in _lib.last_error()
    assert L.mife_assemble_grid(ctypes.byref(desc), ne, 3, 1, 1, 1, None, 1, 3 * rows, 0, 1, None, 1, 1 << 40,
                                None) == 2   # no triquadratic elasticity
    # 8 ranges of ceil(ceil(n_slices / 8) / 4) * 4 entries
    # four int32 per slot, eight ranges of span slots
    assert L.mife_sell_order_len(0) == 0 and L.mife_sell_order_len(1) == 4 * 32
    assert L.mife_sell_order_len(13397) == 4 * 8 * 1676   # ceil(13397 / 8) = 1675 -> 1676


def test_sell3_descriptors_without_gpu():
    """The node-block (SELL3) storage: its format tag and the box builder's block
    fields agree between the header and the bindings; malformed SELL3
    descriptors are refused before any device work; the box sizes report the
    node-block layout for three components only."""
    import re
    from iron_amd import _lib
    text = (ROOT / "include" / "mife.h").read_text()
    assert int(re.search(r"#define MIFE_FMT_SELL3 (\d+)", text).group(1)) == _lib.MIFE_FMT_SELL3
    assert "block" in [f[0] for f in _lib.BoxDesc._fields_]
    assert "block" in [f[0] for f in _lib.BoxCounts._fields_]
    L = _lib.lib()
    M = _lib.MatrixDesc
    F = _lib.MIFE_FMT_SELL3
    # rows not a multiple of three
    d = M(4, F, 1, 1, 1, 1, 36, 1, None, 0)
    assert L.mife_spmv(ctypes.byref(d), 1, 1, 1, 0, None) == 2 and "SELL3" in _lib.last_error()
    # n_slices must be ceil(n / 192)
    d = M(6, F, 1, 1, 1, 2, 36, 1, None, 0)
    assert L.mife_cg(ctypes.byref(d), 1, 1, 1, 0, 1e-8, 10, 1, 1 << 30, None, None, None) == 2
    assert "SELL3" in _lib.last_error()
    # 16-bit node columns are required
    d = M(6, F, 1, 1, 1, 1, 36, None, None, 0)
    assert L.mife_spmv(ctypes.byref(d), 1, 1, 1, 0, None) == 2
    # box sizes: an elasticity box clamped at x = 0 asks for blocks and gets them
    c = _lib.BoxCounts()
    _lib.call("mife_box_sizes", ctypes.byref(_lib.BoxDesc((4, 3, 3), 1, 3, 1, 0, 1, (1.0, 1.0, 1.0), 0.0, 0.0)),
              ctypes.byref(c))
    free_nodes = 4 * 4 * 4   # x = 0 plane fixed: 4 x-planes of 4 x 4 nodes
    assert c.block == 1 and c.col16 == 1 and c.n_free == 3 * free_nodes
    assert c.n_slices == (free_nodes + 63) // 64 and c.stored == 9 * 64 * c.n_slices * c.width
    # a scalar box ignores the request
    _lib.call("mife_box_sizes", ctypes.byref(_lib.BoxDesc((4, 3, 3), 1, 1, 1, 1, 1, (1.0, 1.0, 1.0), 0.0, 1.0)),
              ctypes.byref(c))
    assert c.block == 0


def test_gmres_refinement_mode_checked_without_gpu():
    """The GMRES Gram-Schmidt refinement setter takes 0 (never, the default) or 1
    (always) and refuses anything else."""
    from iron_amd import _lib
    L = _lib.lib()
    assert L.mife_gmres_cgs_refinement(2) == 2 and "refinement" in _lib.last_error()
    assert L.mife_gmres_cgs_refinement(-1) == 2
    assert L.mife_gmres_cgs_refinement(1) == 0
    assert L.mife_gmres_cgs_refinement(0) == 0


def test_one_hip_runtime_in_the_process():
    """torch bundles a HIP runtime with the soname the library links against; the
    binding loads torch's first so that the process holds one runtime (two runtimes
    would make torch's device memory foreign to the library's kernels)."""
    from iron_amd import _lib
    _lib.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(paths) == 1, paths
