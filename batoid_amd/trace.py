"""Per-surface trace functions (batoid/trace.py) on the GPU.

Each call computes the CoordTransform from the rays' frame to the target
frame on the host (coordTransform.py:16-20) and launches one fused HIP
kernel (transform + intersect + interaction [+ obscuration]) over the rays,
in place; the RayVector then reports the target frame."""
from . import _lib, _device
from .coordTransform import CoordTransform


def _niter():
    from .global_vars import _batoid_niter
    return _batoid_niter


def _no_coating(coating):
    if coating is not None:
        raise NotImplementedError("coatings are not part of the GPU trace path")


def applyForwardTransform(ct, rv):
    _lib.call("bt_apply_forward_transform", _lib.dvec(ct.dr), _lib.dvec(ct.drot.ravel()),
              rv._rays, _device.stream())
    rv.coordSys = ct.toSys
    return rv


def applyReverseTransform(ct, rv):
    _lib.call("bt_apply_reverse_transform", _lib.dvec(ct.dr), _lib.dvec(ct.drot.ravel()),
              rv._rays, _device.stream())
    rv.coordSys = ct.fromSys
    return rv


def applyForwardTransformArrays(ct, x, y, z):
    """In-place forward transform of device coordinate columns."""
    _lib.call("bt_apply_forward_transform_arrays", _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), _device.ptr(x), _device.ptr(y),
              _device.ptr(z), x.numel(), _device.stream())


def applyReverseTransformArrays(ct, x, y, z):
    """In-place reverse transform of device coordinate columns."""
    _lib.call("bt_apply_reverse_transform_arrays", _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), _device.ptr(x), _device.ptr(y),
              _device.ptr(z), x.numel(), _device.stream())


def obscure(obsc, rv):
    """Vignette rays where ``obsc`` contains (x, y) (batoid.cpp:151-172)."""
    dev = obsc._dev
    _lib.call("bt_obscure", dev.ptr(), dev.count, rv._rays, _device.stream())
    return rv


def _obsc_args(obscuration):
    if obscuration is None:
        return None, 0
    dev = obscuration._dev
    return dev.ptr(), dev.count


def intersect(surface, rv, coordSys=None, coating=None, obscuration=None):
    """Propagate rays to ``surface`` (in ``coordSys``) in place
    (trace.py:47-81).  ``obscuration``, if given, is applied in the same pass."""
    _no_coating(coating)
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    optr, olen = _obsc_args(obscuration)
    _lib.call("bt_intersect", surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), rv._rays, _niter(), optr, olen,
              _device.stream())
    rv.coordSys = coordSys
    return rv


def reflect(surface, rv, coordSys=None, coating=None, obscuration=None):
    """Intersect and reflect rays in place (trace.py:84-98)."""
    _no_coating(coating)
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    optr, olen = _obsc_args(obscuration)
    _lib.call("bt_reflect", surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), rv._rays, _niter(), optr, olen,
              _device.stream())
    rv.coordSys = coordSys
    return rv


def refract(surface, rv, m1, m2, coordSys=None, coating=None, obscuration=None):
    """Intersect and refract rays in place from ``m1`` into ``m2``
    (trace.py:101-116).  As in the reference kernel, the incoming index is
    taken from |v|, so ``m1`` is not evaluated."""
    _no_coating(coating)
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    optr, olen = _obsc_args(obscuration)
    _lib.call("bt_refract", surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), m2._dev.ptr(), rv._rays, _niter(), optr, olen,
              _device.stream())
    rv.coordSys = coordSys
    return rv


def rSplit(*args, **kwargs):
    raise NotImplementedError("rSplit (ghost splitting) is outside the GPU trace path")


def refractScreen(*args, **kwargs):
    raise NotImplementedError("refractScreen (OPD screens) is outside the GPU trace path")
