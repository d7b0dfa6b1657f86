"""The C-ABI library loads, exports every function include/batoid_amd.h
declares, and agrees with the ctypes binding on struct layouts.  No compute
call is made (none is possible without a GPU)."""
import ctypes
import os
import re

import pytest

from batoid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "batoid_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(bt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("bt_intersect", "bt_reflect", "bt_refract", "bt_obscure",
                     "bt_trace_plan", "bt_sum_amplitude", "bt_huygens_gemm",
                     "bt_huygens_direct", "bt_finish_parallel", "bt_generate_rays"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built: run python -m batoid_amd.build")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares exactly what the header declares
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_struct_layouts_match():
    lib = _lib.load()
    sizes = (ctypes.c_int64 * len(_lib.STRUCTS))()
    assert lib.bt_struct_sizes(sizes, len(_lib.STRUCTS)) == 0
    assert list(sizes) == [ctypes.sizeof(s) for s in _lib.STRUCTS]
    assert lib.bt_abi_version() == 5
