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
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_error_text_roundtrip_without_gpu():
    from iron_amd import _lib
    assert _lib.lib().mife_version().decode().startswith("mi_fe")
    assert _lib.lib().mife_last_error(None, 0) == -1
