"""CPU restatement of the reference's ghost-splitting trace -- test
infrastructure only.

Follows batoid/optic.py:466-535 (Interface.traceSplit with each Interface
kind's rSplit, 786-864) and 1184-1282 (CompoundOptic.traceSplit: the work
queue and the filter of unvignetted rays above minFlux), with every ray
operation run by the C oracle (or_rsplit, or_reflect, or_intersect,
or_obscure) on numpy columns.  Pinned: or_rsplit and the coated
interactions bit-for-bit against the reference's compiled core
(tests/test_oracle_vs_ref.py); the control flow by the reference's own
traceSplit tests, which tests/test_gpu_split.py runs on both.
"""
import ctypes

import numpy as np

import batoid_amd as bt
from . import host as H


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


class Rays:
    """HostRays plus the bookkeeping traceSplit needs (frame, path)."""

    def __init__(self, h, coordSys, path=None):
        self.h = h
        self.coordSys = coordSys
        if path is not None:
            self.path = path

    def copy(self):
        r = Rays(self.h.copy(), self.coordSys)
        if hasattr(self, 'path'):
            r.path = list(self.path)
        return r

    def __len__(self):
        return len(self.h)


def _frame(rays, coordSys):
    ct = bt.CoordTransform(rays.coordSys, coordSys)
    return _dv(ct.dr), _dv(ct.drot)


def _interface_split(item, rays, reverse):
    """rSplit of one Interface kind (optic.py:786-864): (refracted,
    reflected), either possibly None.  Works in place on ``rays``."""
    O = H.load()
    tab = H.host_records(item.surface)
    dr, drot = _frame(rays, item.coordSys)
    kind = type(item).__name__
    if kind == "Mirror":
        O.or_reflect(tab.ptr(), dr, drot, rays.h.struct, 5, 1)
        rays.coordSys = item.coordSys
        return None, rays
    if kind == "Baffle":
        O.or_intersect(tab.ptr(), dr, drot, rays.h.struct, 5, 1)
        rays.coordSys = item.coordSys
        return rays, None
    if kind == "Detector":
        assert not reverse
        m2, coating = item.outMedium, item.forwardCoating
    elif kind == "RefractiveInterface":
        if reverse:
            m2, coating = item.inMedium, item.forwardCoating
        else:
            m2, coating = item.outMedium, item.reverseCoating
    else:
        raise TypeError(kind)
    split = rays.copy()                             # trace.py:125
    med = H.host_records(m2)
    O.or_rsplit(tab.ptr(), dr, drot, med.ptr(), coating.reflectivity, coating.transmissivity,
                rays.h.struct, split.h.struct, 5, 1)
    rays.coordSys = item.coordSys
    split.coordSys = item.coordSys
    return rays, split


def interface_trace_split(item, rays, reverse=False):
    """optic.py:466-535."""
    if item.skip:
        return ([None], [rays]) if reverse else ([rays], [None])
    refracted, reflected = _interface_split(item, rays, reverse)
    if item.obscuration is not None:
        tab = H.host_records(item.obscuration)
        for r in (refracted, reflected):
            if r is not None:
                H.load().or_obscure(tab.ptr(), tab.count, r.h.struct, 1)
    path = list(getattr(rays, 'path', [])) + [item.name]
    for r in (refracted, reflected):
        if r is not None:
            r.path = list(path)
    fwd, rev = [refracted], [reflected]
    if reverse:
        fwd, rev = rev, fwd
    if type(item).__name__ == "Mirror":
        fwd, rev = rev, fwd
    return fwd, rev


def _keep(r, minFlux):
    w = (r.h["vignetted"] == 0) & (r.h["flux"] >= minFlux)
    h = H.HostRays(**{k: v[w] for k, v in r.h.cols.items()})
    return Rays(h, r.coordSys, list(r.path))


def trace_split(optic, rays, minFlux=1e-3, reverse=False):
    """optic.py:1184-1282 (CompoundOptic) or 466-535 (Interface)."""
    if not hasattr(optic, 'items'):
        return interface_trace_split(optic, rays, reverse)
    if optic.skip:
        return ([None], [rays]) if reverse else ([rays], [None])
    work = [(rays, "forward", 0)] if not reverse else [(rays, "reverse", len(optic.items) - 1)]
    outF, outR = [], []
    while work:
        r, direction, index = work.pop()
        item = optic.items[index]
        fwd, rev = trace_split(item, r, minFlux, direction == "reverse")
        fwd = [_keep(x, minFlux) for x in fwd if x]
        rev = [_keep(x, minFlux) for x in rev if x]
        for x in fwd:
            if len(x) > 0:
                if index == len(optic.items) - 1:
                    outF.append(x)
                else:
                    work.append((x, "forward", index + 1))
        for x in rev:
            if len(x) > 0:
                if index == 0:
                    outR.append(x)
                else:
                    work.append((x, "reverse", index - 1))
    return outF, outR
