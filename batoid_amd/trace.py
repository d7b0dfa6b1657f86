"""Per-surface trace functions (batoid/trace.py) on the GPU.

Each call computes the CoordTransform from the rays' frame to the target
frame on the host (coordTransform.py:16-20) and launches one fused HIP
kernel (transform + intersect + interaction [+ obscuration]) over the rays,
in place; the RayVector then reports the target frame."""
import ctypes

from . import _lib, _device
from .coordTransform import CoordTransform


def _niter():
    from .global_vars import _batoid_niter
    return _batoid_niter


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


def _coating_ptr(coating):
    """Host pointer to a bt_coating record (None: no coating); the record is
    copied into the launch, so a temporary suffices."""
    if coating is None:
        return None
    return ctypes.byref(coating._record)


def _interact_live(action, surface, rv, ct, m2, coating, obscuration, live_min_flux,
                   live_counts):
    """bt_interact_live: the interaction for traceSplit's carried sets (rays
    its filter dropped are skipped; the kept rays of the result are counted
    into ``live_counts``, 64 partial sums, if given)."""
    optr, olen = _obsc_args(obscuration)
    kinds = _lib.PLAN_ANALYTIC if surface._analytic else 0
    cptr = _device.ptr(live_counts) if live_counts is not None else None
    _lib.call("bt_interact_live", action, surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), m2._dev.ptr() if m2 is not None else None,
              rv._rays, _niter(), _coating_ptr(coating), optr, olen, float(live_min_flux),
              kinds, cptr, _device.stream())


def intersect(surface, rv, coordSys=None, coating=None, obscuration=None,
              _live_min_flux=None, _live_counts=None):
    """Propagate rays to ``surface`` (in ``coordSys``) in place
    (trace.py:47-81); a coating scales the flux of the rays that reach it by
    its transmissivity.  ``obscuration``, if given, is applied in the same
    pass.  ``_live_min_flux`` / ``_live_counts``: traceSplit's carried sets,
    as for rSplit."""
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    if _live_min_flux is not None:
        _interact_live(_lib.ACT_INTERSECT, surface, rv, ct, None, coating, obscuration,
                       _live_min_flux, _live_counts)
    else:
        optr, olen = _obsc_args(obscuration)
        _lib.call("bt_intersect", surface._dev.ptr(), _lib.dvec(ct.dr),
                  _lib.dvec(ct.drot.ravel()), rv._rays, _niter(), _coating_ptr(coating), optr,
                  olen, _device.stream())
    rv.coordSys = coordSys
    return rv


def reflect(surface, rv, coordSys=None, coating=None, obscuration=None,
            _live_min_flux=None, _live_counts=None):
    """Intersect and reflect rays in place (trace.py:84-98); a coating scales
    the flux by its reflectivity.  ``_live_min_flux`` / ``_live_counts``:
    traceSplit's carried sets, as for rSplit."""
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    if _live_min_flux is not None:
        _interact_live(_lib.ACT_REFLECT, surface, rv, ct, None, coating, obscuration,
                       _live_min_flux, _live_counts)
    else:
        optr, olen = _obsc_args(obscuration)
        _lib.call("bt_reflect", surface._dev.ptr(), _lib.dvec(ct.dr),
                  _lib.dvec(ct.drot.ravel()), rv._rays, _niter(), _coating_ptr(coating), optr,
                  olen, _device.stream())
    rv.coordSys = coordSys
    return rv


def refract(surface, rv, m1, m2, coordSys=None, coating=None, obscuration=None):
    """Intersect and refract rays in place from ``m1`` into ``m2``
    (trace.py:101-116).  As in the reference kernel, the incoming index is
    taken from |v|, so ``m1`` is not evaluated.  A coating scales the flux by
    its transmissivity."""
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    optr, olen = _obsc_args(obscuration)
    _lib.call("bt_refract", surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), m2._dev.ptr(), rv._rays, _niter(),
              _coating_ptr(coating), optr, olen, _device.stream())
    rv.coordSys = coordSys
    return rv


def rSplit(surface, rv, inMedium, outMedium, coating, coordSys=None, obscuration=None,
           _live_min_flux=None, _live_counts=None, _no_split=False):
    """Split rays at ``surface`` into refracted (``rv``, in place) and
    reflected (a new RayVector) parts with the coating's flux fractions
    (trace.py:119-136, batoid.cpp:509-674).  ``obscuration``, if given, is
    applied to both parts in the same pass.  Returns (refracted, reflected).

    ``_live_min_flux`` (traceSplit's own use): ``rv`` still holds rays that
    traceSplit's filter dropped (vignetted, or flux below this threshold);
    they are skipped and only marked dropped in the reflected part, and a
    reflected ray that filter drops is written only as its flux and flags.
    ``_live_counts``: a device int64[2] the numbers of refracted and
    reflected rays passing that filter are added to.  ``_no_split`` (with
    ``_live_min_flux``): traceSplit knows that its filter drops every
    reflected ray; no reflected part is written and None is returned in
    its place."""
    from .rayVector import RayVector
    if coordSys is None:
        coordSys = rv.coordSys
    ct = CoordTransform(rv.coordSys, coordSys)
    no_split = _no_split and _live_min_flux is not None
    split = None if no_split else RayVector._empty(len(rv), coordSys)
    optr, olen = _obsc_args(obscuration)
    args = (surface._dev.ptr(), _lib.dvec(ct.dr), _lib.dvec(ct.drot.ravel()),
            outMedium._dev.ptr(), _coating_ptr(coating), optr, olen, rv._rays,
            _lib.BtRays() if no_split else split._rays)
    if _live_min_flux is None:
        _lib.call("bt_rsplit", *args, _niter(), _device.stream())
    else:
        kinds = _lib.PLAN_ANALYTIC if surface._analytic else 0
        cptr = _device.ptr(_live_counts) if _live_counts is not None else None
        _lib.call("bt_rsplit_live", *args, float(_live_min_flux), _niter(), kinds, cptr,
                  _device.stream())
    rv.coordSys = coordSys
    return rv, split


def refractScreen(*args, **kwargs):
    raise NotImplementedError("refractScreen (OPD screens) is outside the GPU trace path")
