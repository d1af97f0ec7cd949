"""The C-ABI library loads and exports every symbol the public header declares
(CPU only: no compute call is made)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "meeuuw_hip.h")


def header_decls():
    h = open(HEADER).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    out = {}
    for name, params in re.findall(r"^(?:int|long|size_t|const char\*)\s+(mw_\w+)\(([^;]*?)\);", h,
                                   flags=re.M | re.S):
        params = params.strip()
        kinds = []
        if params not in ("void", ""):
            for p in params.split(","):
                p = p.strip()
                kinds.append("P" if "*" in p else "D" if p.startswith("double") else "I")
        out[name] = kinds
    return out


def test_library_exports_header_symbols():
    from meeuuw_amd import _lib
    decls = header_decls()
    assert len(decls) >= 20
    for name in decls:
        assert hasattr(_lib.lib, name), name


def test_ctypes_signatures_match_header():
    from meeuuw_amd import _lib
    decls = header_decls()
    for name, kinds in decls.items():
        res, args = _lib.SIGNATURES[name]
        mine = ["P" if a is _lib.P else "D" if a is _lib.D else "I" for a in args]
        assert mine == kinds, name
    assert set(_lib.SIGNATURES) == set(decls)


def test_no_gpu_needed_for_queries():
    from meeuuw_amd import _lib
    assert _lib.lib.mw_version() >= 100
    assert _lib.lib.mw_sort_workspace_bytes(1000, 100) > 0
    assert _lib.lib.mw_incidence_workspace_bytes(1000) > 8000


def test_errors_map_to_exceptions():
    import pytest
    from meeuuw_amd import _lib
    with pytest.raises(ValueError):
        _lib.call("mw_local_to_global_V", 0, None, None, None)
    assert "local_to_global" in _lib.lib.mw_last_error().decode()


def test_output_argument_errors_without_gpu():
    import ctypes
    import pytest
    from meeuuw_amd import _lib
    h = ctypes.c_void_p()
    with pytest.raises(ValueError):
        _lib.call("mw_output_create", None, b"out", 1, 1, None, ctypes.byref(h))
    assert not h
    with pytest.raises(ValueError):
        _lib.call("mw_output_step", None, None, 0)
    assert _lib.lib.mw_output_destroy(None) == 0
