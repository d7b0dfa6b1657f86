"""ctypes binding of ``libbatoid_amd.so`` (declared in include/batoid_amd.h).

This is the thin host side of the C ABI: the POD descriptor structs, the
loader, and helpers that pack descriptor records into device buffers.  The
library is mandatory: if it is missing, or no GPU is present, every compute
call raises; nothing falls back to CPU arithmetic.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbatoid_amd.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)


class BtRays(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "x", "y", "z", "vx", "vy", "vz", "t", "wavelength", "flux",
        "vignetted", "failed")] + [("n", ctypes.c_int64)]


class BtSurface(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("nchild", ctypes.c_int32),
        ("n1", ctypes.c_int32), ("n2", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("p", ctypes.c_double * 12),
        ("a0", ctypes.c_void_p), ("a1", ctypes.c_void_p),
        ("a2", ctypes.c_void_p), ("a3", ctypes.c_void_p),
    ]


class BtMedium(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("n", ctypes.c_int32),
        ("p", ctypes.c_double * 6),
        ("a0", ctypes.c_void_p), ("a1", ctypes.c_void_p),
    ]


class BtObsc(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32), ("n", ctypes.c_int32),
        ("p", ctypes.c_double * 8),
        ("a0", ctypes.c_void_p), ("a1", ctypes.c_void_p),
    ]


class BtStep(ctypes.Structure):
    _fields_ = [
        ("action", ctypes.c_int32), ("surface", ctypes.c_int32),
        ("medium", ctypes.c_int32), ("obsc", ctypes.c_int32),
        ("obsc_len", ctypes.c_int32), ("start", ctypes.c_int32),
        ("dr", ctypes.c_double * 3), ("drot", ctypes.c_double * 9),
        ("mlp", ctypes.c_void_p),
    ]


class BtPlan(ctypes.Structure):
    _fields_ = [
        ("nstep", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("nmedia", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("steps", ctypes.c_void_p), ("surfaces", ctypes.c_void_p),
        ("media", ctypes.c_void_p), ("obscs", ctypes.c_void_p),
    ]


class BtPsfSegment(ctypes.Structure):
    _fields_ = [
        ("begin", ctypes.c_int64), ("end", ctypes.c_int64),
        ("chief", ctypes.c_int64), ("time_ray", ctypes.c_int64),
        ("a1", ctypes.c_double * 2), ("a2", ctypes.c_double * 2),
    ]


class BtRaygen(ctypes.Structure):
    _fields_ = [
        ("pattern", ctypes.c_int32), ("nring", ctypes.c_int32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double),
        ("lx", ctypes.c_double * 2), ("ly", ctypes.c_double * 2),
        ("lo", ctypes.c_double * 2), ("span", ctypes.c_double * 2),
        ("pcg_state", ctypes.c_uint64 * 2), ("pcg_inc", ctypes.c_uint64 * 2),
        ("ring_off", ctypes.c_void_p), ("ring_rho", ctypes.c_void_p),
        ("ring_dth", ctypes.c_void_p), ("px", ctypes.c_void_p), ("py", ctypes.c_void_p),
        ("flux", ctypes.c_void_p), ("stop", ctypes.c_void_p),
        ("stop_dr", ctypes.c_double * 3), ("stop_drot", ctypes.c_double * 9),
        ("origin", ctypes.c_double * 3), ("rot", ctypes.c_double * 9),
        ("vv", ctypes.c_double * 3), ("wavelength", ctypes.c_double),
        ("flux0", ctypes.c_double),
    ]


class BtCoating(ctypes.Structure):
    _fields_ = [("reflectivity", ctypes.c_double), ("transmissivity", ctypes.c_double)]


STRUCTS = (BtRays, BtSurface, BtMedium, BtObsc, BtStep, BtPlan, BtPsfSegment, BtRaygen,
           BtCoating)

# enum values (include/batoid_amd.h)
PLANE, SPHERE, PARABOLOID, QUADRIC, ASPHERE, TILTED, BICUBIC, POLYNOMIAL, SUM = range(9)
MED_CONST, MED_TABLE, MED_SELLMEIER, MED_SUMITA, MED_AIR = range(5)
(OB_CIRCLE, OB_ANNULUS, OB_RECTANGLE, OB_RAY, OB_POLYGON, OB_NEGATION,
 OB_UNION, OB_INTERSECTION) = range(8)
ACT_INTERSECT, ACT_REFLECT, ACT_REFRACT = range(3)
START_REFERENCE, START_CONIC, START_MLP = range(3)
PLAN_NO_REFRACT = 1
PLAN_FAST_START = 2
PLAN_MLP_START = 4
PLAN_ANALYTIC = 8
PLAN_SUM_APB = 16
PLAN_FLAT_SUMS = 32
ANALYTIC_KINDS = (PLANE, SPHERE, PARABOLOID, QUADRIC, ASPHERE, TILTED)
PSF_FP32, PSF_BF16X3 = 0, 1
GEN_POINTS, GEN_GRID, GEN_GRID_RANDOM, GEN_POLAR, GEN_POLAR_RANDOM = range(5)
OBSC_MAX_STACK = 32
MAX_SUM_DEPTH = 3

# name -> (restype, argtypes)
_V = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_L = ctypes.c_int64
SIGNATURES = {
    "bt_abi_version": (_I, []),
    "bt_last_error": (ctypes.c_char_p, []),
    "bt_struct_sizes": (_I, [ctypes.POINTER(ctypes.c_int64), _I]),
    "bt_struct_field_count": (_I, [_I]),
    "bt_struct_field_offset": (_L, [_I, ctypes.c_char_p]),
    "bt_device_info": (_I, [ctypes.POINTER(_I)] * 3),
    "bt_apply_forward_transform": (_I, [_V, _V, BtRays, _V]),
    "bt_apply_reverse_transform": (_I, [_V, _V, BtRays, _V]),
    "bt_apply_forward_transform_arrays": (_I, [_V, _V, _V, _V, _V, _L, _V]),
    "bt_apply_reverse_transform_arrays": (_I, [_V, _V, _V, _V, _V, _L, _V]),
    "bt_finish_parallel": (_I, [_V, _V, _V, _V, _V, _V, _L, _V]),
    "bt_generate_rays": (_I, [_V, BtRays, _V]),
    "bt_intersect": (_I, [_V, _V, _V, BtRays, _I, _V, _V, _I, _V]),
    "bt_reflect": (_I, [_V, _V, _V, BtRays, _I, _V, _V, _I, _V]),
    "bt_refract": (_I, [_V, _V, _V, _V, BtRays, _I, _V, _V, _I, _V]),
    "bt_rsplit": (_I, [_V, _V, _V, _V, _V, _V, _I, BtRays, BtRays, _I, _V]),
    "bt_rsplit_live": (_I, [_V, _V, _V, _V, _V, _V, _I, BtRays, BtRays, _D, _I, _I, _V, _V]),
    "bt_interact_live": (_I, [_I, _V, _V, _V, _V, BtRays, _I, _V, _V, _I, _D, _I, _V, _V]),
    "bt_select_workspace_bytes": (_L, [_L]),
    "bt_select_count": (_I, [BtRays, _D, _V, _V, _L, _V]),
    "bt_select_scatter": (_I, [BtRays, _D, BtRays, _V, _L, _V]),
    "bt_obscure": (_I, [_V, _I, BtRays, _V]),
    "bt_trace_plan": (_I, [BtPlan, BtRays, BtRays, _I, _V]),
    "bt_surface_sag": (_I, [_V, _V, _V, _V, _L, _V]),
    "bt_surface_normal": (_I, [_V, _V, _V, _V, _L, _V]),
    "bt_medium_get_n": (_I, [_V, _V, _V, _L, _V]),
    "bt_obsc_contains": (_I, [_V, _I, _V, _V, _V, _L, _V]),
    "bt_check_sqrt": (_I, [_V, _V, _L, _V]),
    "bt_position_at_time": (_I, [BtRays, _D, _V, _V, _V, _V]),
    "bt_propagate_in_place": (_I, [BtRays, _D, _V]),
    "bt_phase": (_I, [BtRays, _D, _D, _D, _D, _V, _V]),
    "bt_amplitude": (_I, [BtRays, _D, _D, _D, _D, _V, _V]),
    "bt_sum_amplitude_workspace_bytes": (_L, [_L]),
    "bt_sum_amplitude": (_I, [BtRays, _D, _D, _D, _D, _I, _V, _V, _V]),
    "bt_mlp_param_count": (_I, []),
    "bt_mlp_packed_bytes": (_L, []),
    "bt_mlp_grad_workspace_bytes": (_L, [_L]),
    "bt_mlp_loss_grad": (_I, [_V, _V, _V, _V, _L, _V, _V, _V, _V, _L, _V]),
    "bt_mlp_adam": (_I, [_V, _V, _V, _V, _L, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, ctypes.c_float, _I, _V]),
    "bt_mlp_forward": (_I, [_V, _V, _V, _L, _V, _V]),
    "bt_mlp_pack": (_I, [_V, _V, _V, _V]),
    "bt_mlp_forward_packed": (_I, [_V, _V, _L, _V, _V]),
    "bt_newton_samples": (_I, [_V, _V, _V, BtRays, _I, _V, _V, _V, _V, _V]),
    "bt_huygens_workspace_bytes": (_L, [_L, _I, _I]),
    "bt_huygens_direct": (_I, [BtRays, _V, _V, _V, _I, _I, _D, _I, _V, _V, _L, _V]),
    "bt_huygens_gemm": (_I, [BtRays, _V, _V, _V, _I, _I, _D, _I, _I, _V, _V, _L, _V]),
    "bt_huygens_noalloc_scratch_bytes": (_L, []),
    "bt_huygens_gemm_noalloc": (_I, [BtRays, _V, _V, _V, _I, _I, _D, _I, _I, _V, _V]),
    "bt_huygens_batch_workspace_bytes": (_L, [_L, _I, _I, _I, _I]),
    "bt_huygens_batch": (_I, [BtRays, _V, _I, _I, _I, _I, _I, _I, _I, _V, _V, _L, _V]),
    "bt_zernike_workspace_bytes": (_L, [_I]),
    "bt_zernike_gram": (_I, [_V, _V, _V, _V, _V, _L, _D, _D, _D, _V, _I, _I, _V, _V, _L, _V]),
}


class BatoidAMDError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (once) and declare every C-ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BatoidAMDError(
            f"{path} is missing: build it with `python -m batoid_amd.build`"
        )
    lib = ctypes.CDLL(path)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    sizes = (ctypes.c_int64 * len(STRUCTS))()
    lib.bt_struct_sizes(sizes, len(STRUCTS))
    for st, sz in zip(STRUCTS, sizes):
        if ctypes.sizeof(st) != sz:
            raise BatoidAMDError(
                f"ABI mismatch: sizeof({st.__name__}) is {sz} in the library, "
                f"{ctypes.sizeof(st)} in the binding"
            )
    for which, st in enumerate(STRUCTS):
        mismatch = layout_mismatch(lib, which, st)
        if mismatch:
            raise BatoidAMDError(f"ABI mismatch in {st.__name__}: {mismatch}")
    _lib = lib
    return lib


def layout_mismatch(lib, which, st):
    """'' if every field of ctypes struct ``st`` sits, by name, at the offset
    the library reports for struct number ``which``, else a description."""
    names = [f[0] for f in st._fields_]
    count = lib.bt_struct_field_count(which)
    if count != len(names):
        return f"{count} fields in the library, {len(names)} in the binding"
    bad = []
    for n in names:
        off = lib.bt_struct_field_offset(which, n.encode())
        if off != getattr(st, n).offset:
            bad.append(f"{n} at {off} (binding {getattr(st, n).offset})")
    return "; ".join(bad)


def check(rc, what=""):
    if rc != 0:
        msg = load().bt_last_error()
        msg = msg.decode() if msg else ""
        raise BatoidAMDError(f"{what} failed with status {rc}: {msg}")


def call(name, *args):
    """Call a C-ABI function and raise on a nonzero status."""
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


def dvec(values):
    """A host double array argument (dr, drot, ...)."""
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr


def struct_array_bytes(records):
    """Raw bytes of a ctypes struct sequence, for upload."""
    if not records:
        return b""
    typ = type(records[0])
    arr = (typ * len(records))(*records)
    return bytes(memoryview(arr).cast("B"))


def np_ptr(a):
    return a.ctypes.data if a is not None else None
