"""The C-ABI library loads and exports every entry point include/dfx.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

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


def test_serving_footprint_without_device():
    """The footprint estimate is host arithmetic: larger meshes and more states
    cost more, and a request larger than the budget is refused up front."""
    import dfx
    base = dict(p=3, n=(6, 6, 6), lo=(-8.0,) * 3, hi=(8.0,) * 3, atoms=[(0.0, 0.0, 0.0, 6)],
                enrich=True, states=10, electrons=6.0)
    f0 = dfx.request_footprint(base)
    assert f0 > 0
    assert dfx.request_footprint(dict(base, n=(10, 10, 10))) > f0
    assert dfx.request_footprint(dict(base, states=40)) > f0
    with pytest.raises(dfx.DfxError):
        dfx.serve([base], max_slots=1, memory_budget=f0 // 2)
