"""The C-ABI library loads and exports every entry point include/dfx.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfx_[a-z0-9_]+)\s*\(", text)))


def test_libdfx_exports_every_declared_symbol():
    import dfx
    lib = ctypes.CDLL(dfx.LIB_PATH)
    names = declared("dfx.h")
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_binds_only_declared_symbols():
    import dfx
    names = set(declared("dfx.h"))
    bound = set(dfx.exported_symbols())
    assert bound <= names, sorted(bound - names)


def test_error_paths_without_device():
    """Argument errors are reported as negative codes without touching a GPU."""
    import dfx
    lib = dfx._lib
    assert lib.dfx_mv_destroy(None) == -1
    assert lib.dfx_op_apply(None, None, None) == -1
    assert lib.dfx_chfsi_timings(None, None) == -1
