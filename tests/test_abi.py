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


def test_struct_field_offsets_match():
    """Every field of every struct sits, by name, where the library puts it
    (two same-size fields swapped in the binding would pass the size check)."""
    lib = _lib.load()
    for which, st in enumerate(_lib.STRUCTS):
        assert lib.bt_struct_field_count(which) == len(st._fields_), st.__name__
        assert _lib.layout_mismatch(lib, which, st) == "", st.__name__
    assert lib.bt_struct_field_count(len(_lib.STRUCTS)) < 0
    assert lib.bt_struct_field_offset(0, b"no_such_field") == -1


def test_offset_check_catches_swapped_fields():
    """A binding with two 8-byte fields swapped is rejected, field by field."""
    lib = _lib.load()

    class Reordered(ctypes.Structure):
        _fields_ = [("transmissivity", ctypes.c_double), ("reflectivity", ctypes.c_double)]
    assert ctypes.sizeof(Reordered) == ctypes.sizeof(_lib.BtCoating)
    msg = _lib.layout_mismatch(lib, _lib.STRUCTS.index(_lib.BtCoating), Reordered)
    assert "transmissivity" in msg and "reflectivity" in msg


def test_noalloc_gemm_refuses_oversized_requests_without_a_gpu():
    """The allocation-free PSF entry point checks its fixed scratch on the
    host: a request beyond it is refused with BT_E_SIZE before any HIP call
    (the pointers below are never dereferenced)."""
    lib = _lib.load()
    cap = lib.bt_huygens_noalloc_scratch_bytes()
    r = _lib.BtRays()
    for k, _ in _lib.BtRays._fields_[:-1]:
        setattr(r, k, 0x1000)
    r.n = 1 << 22
    d = (ctypes.c_double * 2)(0.0, 0.0)
    rc = lib.bt_huygens_gemm_noalloc(r, d, d, d, 64, 64, 0.0, 1, _lib.PSF_FP32,
                                     ctypes.c_void_p(0x1000), None)
    assert rc == -2
    msg = lib.bt_last_error().decode()
    assert "scratch" in msg and str(cap) in msg
    # and null pointers are refused first
    assert lib.bt_huygens_gemm_noalloc(_lib.BtRays(), d, d, d, 64, 64, 0.0, 1,
                                       _lib.PSF_FP32, None, None) == -1
