"""Optical elements and systems (batoid/optic.py).

``Interface`` subclasses choose the interaction (intersect / reflect /
refract); ``CompoundOptic.trace`` compiles its unskipped Interfaces into a
plan and traces every ray through the whole optic in one fused GPU kernel
(per-ray results identical to calling each Interface in turn)."""
from functools import cached_property

import numpy as np
import torch

from . import _lib, _device
from .constants import globalCoordSys, vacuum
from .obscuration import ObscNegation, ObscCircle, ObscAnnulus
from .coating import SimpleCoating
from .rayVector import RayVector


_DERIVED = ('path', 'itemDict', '_names', 'R_inner', 'R_outer', 'inRadius',
            'outRadius', '_plan_cache')


def _init_kwargs(obj, drop=()):
    """Constructor keywords that rebuild ``obj`` (minus derived state)."""
    return {k: v for k, v in obj.__dict__.items()
            if k not in _DERIVED and k not in drop}


class Optic:
    """Base class: name, frame, media, skip flag, plus any extra attributes
    (backDist, pupilSize, stopSurface, ...) as keyword arguments
    (optic.py:13-47)."""

    def __init__(self, name=None, coordSys=globalCoordSys, inMedium=vacuum,
                 outMedium=vacuum, skip=False, **kwargs):
        self.name = name
        self.coordSys = coordSys
        self.inMedium = inMedium
        self.outMedium = outMedium
        self.skip = False
        kwargs.pop('itemDict', None)
        self.__dict__.update(**kwargs)

    @classmethod
    def fromYaml(cls, filename):
        """Load an optic from a batoid YAML description (optic.py:59-95)."""
        from .parse import load_yaml, parse_optic
        return parse_optic(load_yaml(filename)['opticalSystem'])

    # ---- rigid-body perturbations (optic.py:571-697, 1362-1566) ----------
    def _moved(self, frame, stop, sub):
        """Copy with coordSys -> frame(coordSys), stopSurface ->
        stop(stopSurface) and, for compound optics, each item -> sub(item)."""
        kw = _init_kwargs(self, (self._lead,))
        kw['coordSys'] = frame(self.coordSys)
        if kw.get('stopSurface') is not None:
            kw['stopSurface'] = stop(kw['stopSurface'])
        lead = getattr(self, self._lead)
        if self._lead == 'items':
            lead = [sub(item) for item in lead]
        ret = self.__class__.__new__(self.__class__)
        ret.__init__(lead, **kw)
        return ret

    def withGlobalShift(self, shift):
        """Copy shifted by ``shift`` (global axes), stop surface and
        sub-items included."""
        shift = np.asarray(shift, dtype=float)
        again = lambda o: o.withGlobalShift(shift)
        return self._moved(lambda cs: cs.shiftGlobal(shift), again, again)

    def withLocalShift(self, shift):
        """Copy shifted by ``shift`` along this optic's own axes."""
        return self.withGlobalShift(self.coordSys.rot @ np.asarray(shift, dtype=float))

    def withGlobalRotation(self, rot, rotCenter=None, coordSys=None):
        """Copy rotated by ``rot`` (global axes) about ``rotCenter`` given in
        ``coordSys`` (default global; 'local' = this optic's frame)."""
        center = [0, 0, 0] if rotCenter is None else rotCenter
        if coordSys is None:
            cs = globalCoordSys
        elif isinstance(coordSys, str):
            if coordSys != 'local':
                raise ValueError(f"unknown coordSys {coordSys!r}")
            cs = self.coordSys
        else:
            cs = coordSys
        again = lambda o: o.withGlobalRotation(rot, center, cs)
        return self._moved(lambda c: c.rotateGlobal(rot, center, cs), again, again)

    def withLocalRotation(self, rot, rotCenter=None, coordSys=None):
        """Copy rotated by ``rot`` about this optic's own axes, about
        ``rotCenter`` given in ``coordSys`` (default this optic's frame)."""
        center = [0, 0, 0] if rotCenter is None else rotCenter
        cs = self.coordSys if coordSys is None else coordSys
        again = lambda o: o.withLocalRotation(rot, center, cs)
        return self._moved(lambda c: c.rotateLocal(rot, center, cs), again, again)

    def _repr_helper(self):
        out = ""
        if self.name is not None:
            out += ", name={!r}".format(self.name)
        out += ", coordSys={!r}".format(self.coordSys)
        if self.inMedium != vacuum:
            out += ", inMedium={!r}".format(self.inMedium)
        if self.outMedium != vacuum:
            out += ", outMedium={!r}".format(self.outMedium)
        return out


class Interface(Optic):
    """A single surface with an optional obscuration; rays pass through
    (optic.py:98-140, 541-543)."""
    _lead = 'surface'

    def __init__(self, surface, obscuration=None, **kwargs):
        Optic.__init__(self, **kwargs)
        self.path = [self.name]
        self.surface = surface
        self.obscuration = obscuration
        self.inRadius = self.R_inner
        self.outRadius = self.R_outer

    def trace(self, rv, reverse=False):
        """Interact and then obscure, in place (optic.py:391-424); both run in
        one GPU pass."""
        self._interact(rv, reverse, self.obscuration)
        return rv

    def _interact(self, rv, reverse, obscuration):
        from .trace import intersect
        return intersect(self.surface, rv, self.coordSys, obscuration=obscuration)

    def interact(self, rv, reverse=False):
        return self._interact(rv, reverse, None)

    def intersect(self, rv, reverse=False):
        from .trace import intersect
        return intersect(self.surface, rv, self.coordSys)

    def traceFull(self, rv, reverse=False):
        """{name: {'name', 'in', 'out'}} for this single Interface."""
        from collections import OrderedDict
        out = OrderedDict()
        rin = rv.copy()
        self.trace(rv, reverse=reverse)
        out[self.name] = {'name': self.name, 'in': rin, 'out': rv.copy()}
        return out

    def clearObscuration(self, unless=()):
        if self.name not in unless:
            self.obscuration = None

    def traceSplit(self, rv, minFlux=1e-3, reverse=False, _verbose=False, _filtered=False,
                   _counts=None):
        """Split rays at this Interface into those continuing forward and
        those sent back (optic.py:466-535); returns two lists of
        RayVectors (None entries where a direction gets no rays), each
        tagged with its ``path``.

        ``_filtered`` (CompoundOptic.traceSplit's own use): ``rv`` still
        holds the rays traceSplit's filter dropped after the previous item
        (vignetted, or flux < minFlux); they are carried along, untouched,
        instead of being copied out.  ``_counts``: (device int64 array,
        index) of two counts the split adds the numbers of refracted and
        reflected rays passing the filter to, each spread over
        ``_COUNT_SLOTS`` entries (the outputs' ``_live_slot``)."""
        if _verbose:
            s = "forward" if not reverse else "reverse"
            print("traceSplit {}       {:30s} flux = {:18.8f}   nphot = {:10d}".format(
                s, self.name, *_live_stats(rv, minFlux if _filtered else None)))
        if self.skip:
            rv._live_slot = rv._live_count = None
            return (None, rv) if reverse else (rv, None)
        # the obscuration applies to both directions, in the same pass
        live = _filtered and isinstance(self, (RefractiveInterface, Detector, Mirror, Baffle))
        live_counts = None
        if live and _counts is not None:
            base, off = _counts
            live_counts = base[off * _COUNT_SLOTS:(off + 2) * _COUNT_SLOTS]
        bound = getattr(rv, '_flux_bound', None)
        no_split = False
        if bound is not None:
            ft, fr = self._split_fractions(reverse)
            # the filter drops every reflected ray: that part is not written
            no_split = live and not bound * abs(fr) >= minFlux
        refracted, reflected = self.rSplit(
            rv, reverse=reverse, obscuration=self.obscuration,
            _live_min_flux=minFlux if _filtered else None, _live_counts=live_counts,
            _no_split=no_split)
        if bound is not None:
            if refracted is not None:
                refracted._flux_bound = bound * abs(ft)
            if reflected is not None:
                reflected._flux_bound = bound * abs(fr)
        path = list(getattr(rv, 'path', [])) + [self.name]
        for k, r in enumerate((refracted, reflected)):
            if r is not None:
                r.path = list(path)
                r._live_slot = (base, off + k) if live_counts is not None else None
                r._live_count = None
        rForward, rReverse = [refracted], [reflected]
        if reverse:
            rForward, rReverse = rReverse, rForward
        if isinstance(self, Mirror):
            rForward, rReverse = rReverse, rForward
        return rForward, rReverse

    def _split_batch(self, inputs, minFlux, _verbose=False, _compact=True, _deferred=False):
        """traceSplit of several (rays, reverse, filtered) inputs (no host
        synchronisation: the launches queue on the stream); the filtered
        inputs' splits count their kept rays into one device array (not
        when ``_deferred``: nobody reads them)."""
        counts = None
        if not _deferred and any(filt for _, _, filt in inputs):
            counts = torch.zeros(2 * _COUNT_SLOTS * len(inputs), dtype=torch.int64,
                                 device=_device.device())
        return [self.traceSplit(rv, minFlux, rev, _verbose, filt,
                                (counts, 2 * j) if counts is not None else None)
                for j, (rv, rev, filt) in enumerate(inputs)]

    def _split_fractions(self, reverse):
        """Factors rSplit applies to the flux of its (refracted, reflected)
        parts: 1 for a step without a coating (Mirror, Baffle)."""
        return 1.0, 1.0

    @cached_property
    def R_outer(self):
        obsc = self.obscuration
        if obsc is not None:
            if isinstance(obsc, ObscNegation):
                obsc = obsc.original
            if isinstance(obsc, ObscAnnulus):
                return obsc.outer
            if isinstance(obsc, ObscCircle):
                return obsc.radius
        return None

    @cached_property
    def R_inner(self):
        obsc = self.obscuration
        if obsc is not None:
            if isinstance(obsc, ObscNegation):
                obsc = obsc.original
            if isinstance(obsc, ObscAnnulus):
                return obsc.inner
            if isinstance(obsc, ObscCircle):
                return 0.0
        return None

    def withSurface(self, surface):
        """Copy of this Interface with ``surface`` replaced (optic.py:698-715)."""
        ret = self.__class__.__new__(self.__class__)
        kw = _init_kwargs(self, ('surface',))
        ret.__init__(surface, **kw)
        return ret

    def withPerturbedSurface(self, perturbation):
        """Copy with ``surface + perturbation`` (optic.py:717-731)."""
        return self.withSurface(self.surface + perturbation)

    def __hash__(self):
        return hash((self.__class__.__name__, self.surface, self.obscuration,
                     self.name, self.inMedium, self.outMedium, self.coordSys))

    def __eq__(self, other):
        if not self.__class__ == other.__class__:
            return False
        return (self.surface == other.surface and self.obscuration == other.obscuration
                and self.name == other.name and self.inMedium == other.inMedium
                and self.outMedium == other.outMedium and self.coordSys == other.coordSys)

    def __ne__(self, other):
        return not (self == other)

    def __repr__(self):
        out = "{!s}({!r}".format(self.__class__.__name__, self.surface)
        if self.obscuration is not None:
            out += ", obscuration={!r}".format(self.obscuration)
        return out + Optic._repr_helper(self) + ")"


class RefractiveInterface(Interface):
    """Refracts from inMedium into outMedium (reverse swaps them)
    (optic.py:762-797).  Coatings default to full transmission."""

    def __init__(self, *args, **kwargs):
        Interface.__init__(self, *args, **kwargs)
        self.forwardCoating = SimpleCoating(reflectivity=0.0, transmissivity=1.0)
        self.reverseCoating = SimpleCoating(reflectivity=0.0, transmissivity=1.0)

    def rSplit(self, rv, reverse=False, obscuration=None, _live_min_flux=None,
               _live_counts=None, _no_split=False):
        """(refracted, reflected); as optic.py:786-797, the forward direction
        uses reverseCoating and the reverse direction forwardCoating."""
        if reverse:
            m1, m2, coating = self.outMedium, self.inMedium, self.forwardCoating
        else:
            m1, m2, coating = self.inMedium, self.outMedium, self.reverseCoating
        return self.surface.rSplit(rv, m1, m2, coating, coordSys=self.coordSys,
                                   obscuration=obscuration, _live_min_flux=_live_min_flux,
                                   _live_counts=_live_counts, _no_split=_no_split)

    def _split_fractions(self, reverse):
        coating = self.forwardCoating if reverse else self.reverseCoating
        return coating.transmissivity, coating.reflectivity

    def _interact(self, rv, reverse, obscuration):
        from .trace import refract
        m1, m2 = (self.outMedium, self.inMedium) if reverse else (self.inMedium, self.outMedium)
        return refract(self.surface, rv, m1, m2, self.coordSys, obscuration=obscuration)

    def refract(self, rv, reverse=False):
        return self.interact(rv, reverse=reverse)


class Mirror(Interface):
    """Reflects (optic.py:800-823)."""

    def __init__(self, *args, **kwargs):
        Interface.__init__(self, *args, **kwargs)
        self.forwardCoating = SimpleCoating(reflectivity=1.0, transmissivity=0.0)
        self.reverseCoating = SimpleCoating(reflectivity=1.0, transmissivity=0.0)

    def rSplit(self, rv, reverse=False, obscuration=None, _live_min_flux=None,
               _live_counts=None, _no_split=False):
        from .trace import reflect
        counts = _live_counts[_COUNT_SLOTS:] if _live_counts is not None else None
        return None, reflect(self.surface, rv, self.coordSys, obscuration=obscuration,
                             _live_min_flux=_live_min_flux, _live_counts=counts)

    def _interact(self, rv, reverse, obscuration):
        from .trace import reflect
        return reflect(self.surface, rv, self.coordSys, obscuration=obscuration)

    def reflect(self, rv, reverse=False):
        return self.interact(rv, reverse=reverse)


class Detector(Interface):
    """Rays stop here (intersect only) (optic.py:826-845)."""

    def __init__(self, *args, **kwargs):
        Interface.__init__(self, *args, **kwargs)
        self.forwardCoating = SimpleCoating(reflectivity=0.0, transmissivity=1.0)
        self.reverseCoating = None

    def rSplit(self, rv, reverse=False, obscuration=None, _live_min_flux=None,
               _live_counts=None, _no_split=False):
        assert not reverse
        return self.surface.rSplit(rv, self.inMedium, self.outMedium, self.forwardCoating,
                                   coordSys=self.coordSys, obscuration=obscuration,
                                   _live_min_flux=_live_min_flux, _live_counts=_live_counts,
                                   _no_split=_no_split)

    def _split_fractions(self, reverse):
        return self.forwardCoating.transmissivity, self.forwardCoating.reflectivity


class Baffle(Interface):
    """Rays pass straight through; only vignetting applies (optic.py:848-864)."""

    def __init__(self, *args, **kwargs):
        Interface.__init__(self, *args, **kwargs)
        self.forwardCoating = SimpleCoating(reflectivity=0.0, transmissivity=1.0)
        self.reverseCoating = SimpleCoating(reflectivity=0.0, transmissivity=1.0)

    def rSplit(self, rv, reverse=False, obscuration=None, _live_min_flux=None,
               _live_counts=None, _no_split=False):
        from .trace import intersect
        counts = _live_counts[:_COUNT_SLOTS] if _live_counts is not None else None
        return intersect(self.surface, rv, self.coordSys, obscuration=obscuration,
                         _live_min_flux=_live_min_flux, _live_counts=counts), None


class OPDScreen(Interface):
    """Surface plus an optical-path-difference screen (optic.py:867-930).
    Tracing through it (refractScreen) is outside the GPU trace path."""

    def __init__(self, surface, screen, **kwargs):
        Interface.__init__(self, surface, **kwargs)
        self.screen = screen

    def _interact(self, rv, reverse, obscuration):
        raise NotImplementedError("OPDScreen is outside the GPU trace path")

    def rSplit(self, rv, reverse=False, obscuration=None, _live_min_flux=None,
               _live_counts=None, _no_split=False):
        raise NotImplementedError("OPDScreen is outside the GPU trace path")


# partial sums per count that bt_rsplit_live spreads its atomics over
_COUNT_SLOTS = 64


def _count_children(children, minFlux, compact):
    """The number of rays of each child set that pass traceSplit's filter,
    read back in one synchronisation: from the split kernel's own counts
    (``_live_slot``), from a nested optic (``_live_count``), or counted by
    bt_select_count -- always so for sets that are to be copied out
    (leaving the optic with ``compact``), whose copy needs that pass's
    offsets.  Returns (counts, copy), copy(i, m) giving child i's kept rays
    as a new RayVector."""
    need, slot_of = [], {}
    bases = []
    for ci, (_, _, rr, leaves) in enumerate(children):
        if leaves and compact:
            need.append(ci)
        elif getattr(rr, '_live_count', None) is not None:
            continue
        elif getattr(rr, '_live_slot', None) is not None:
            base = rr._live_slot[0]
            if id(base) not in slot_of:
                slot_of[id(base)] = len(bases)
                bases.append(base)
        else:
            need.append(ci)
    sel = RayVector._count_select([children[ci][2] for ci in need], minFlux) if need else None
    parts = [b.view(-1, _COUNT_SLOTS).sum(1) for b in bases]
    parts += [sel.counts] if sel is not None else []
    vals = torch.cat(parts).tolist() if parts else []
    starts, pos = [], 0
    for b in bases:
        starts.append(pos)
        pos += b.numel() // _COUNT_SLOTS
    need_at = {ci: j for j, ci in enumerate(need)}
    counts = []
    for ci, (_, _, rr, leaves) in enumerate(children):
        if ci in need_at:
            counts.append(int(vals[pos + need_at[ci]]))
        elif getattr(rr, '_live_count', None) is not None:
            counts.append(int(rr._live_count))
        else:
            base, off = rr._live_slot
            counts.append(int(vals[starts[slot_of[id(base)]] + off]))
    return counts, (lambda ci, m: sel.scatter(need_at[ci], m))


def _live_stats(rv, minFlux):
    """(flux sum, ray count) of the rays of ``rv`` traceSplit still carries:
    all of them, or with ``minFlux`` those unvignetted with flux >= minFlux."""
    flux = rv.torch("flux")
    if minFlux is None:
        return float(flux.sum()), len(rv)
    keep = (rv.torch("vignetted") == 0) & (flux >= minFlux)
    return float(flux[keep].sum()), int(keep.sum())


class CompoundOptic(Optic):
    """A sequence of optics traced in order (optic.py:933-1083)."""
    _lead = 'items'

    def __init__(self, items, **kwargs):
        Optic.__init__(self, **kwargs)
        self.items = tuple(items)
        self.path = [name for item in self.items for name in item.path]

    @cached_property
    def itemDict(self):
        out = {self.name: self}
        for item in self.items:
            out[self.name + '.' + item.name] = item
            if hasattr(item, 'itemDict'):
                for k, v in item.itemDict.items():
                    out[self.name + '.' + k] = v
        return out

    @cached_property
    def _names(self):
        nameDict = {}
        duplicates = set()
        for k in self.itemDict.keys():
            tokens = k.split('.')
            shortNames = [tokens[-1]]
            for token in reversed(tokens[:-1]):
                shortNames.append('.'.join([token, shortNames[-1]]))
            for shortName in shortNames:
                if shortName in nameDict:
                    duplicates.add(shortName)
                else:
                    nameDict[shortName] = k
        for name in duplicates:
            del nameDict[name]
        return nameDict

    def __getitem__(self, key):
        try:
            return self.itemDict[key]
        except KeyError:
            try:
                return self.itemDict[self._names[key]]
            except KeyError:
                raise ValueError("Cannot find item {}".format(key))

    def _skips(self):
        flags = []

        def visit(o):
            flags.append(o.skip)
            for it in getattr(o, 'items', ()):
                visit(it)
        visit(self)
        return tuple(flags)

    def _plan(self, coordSys_in, reverse, start=None):
        """The device plan of the optic as it is now.  Items' surfaces,
        obscurations, media and frames may be reassigned between traces
        (``optic['M1'].obscuration = ...``), so the cache key names every
        step's objects; the entry keeps those objects alive, so their ids
        cannot be reused while the entry exists."""
        from .plan import PlanSpec, DevicePlan, collect_steps, start_key
        steps = collect_steps(self, reverse)
        ident = tuple((id(i), id(i.surface), id(i.obscuration), id(m2), i.coordSys, action)
                      for i, action, m2 in steps)
        key = (coordSys_in, reverse, ident, start_key(start))
        cache = self.__dict__.setdefault('_plan_cache', {})
        entry = cache.get(key)
        if entry is None:
            if len(cache) >= 16:                     # bounded: drop the oldest
                cache.pop(next(iter(cache)))
            entry = (DevicePlan(PlanSpec(steps, coordSys_in, start)), steps)
            cache[key] = entry
        return entry[0]

    def trace(self, rv, reverse=False, path=None, start=None):
        """Trace rays through every unskipped item, in place
        (optic.py:1013-1083).  Without ``path`` the whole optic runs as one
        fused GPU kernel; with ``path`` the items are traced one by one in
        the requested order.

        ``start`` picks how Newton surfaces seed their intersection:
        "reference" (default; bit-identical to the reference), "conic" (an
        analytic start -- the base conic's root, one Newton update on the
        leading Asphere of a Sum, or a Sum's first child's reference root --
        converged to a 1e-14 m sag gap), or a mapping
        {interface name: NewtonStartNet} (learned starts, see
        newton_start.py).  Fast starts fall back to the reference path for
        any ray they do not converge."""
        from .global_vars import _batoid_niter
        if path is not None:
            if start not in (None, "reference"):
                raise NotImplementedError("fast starts apply to the fused trace only")
            return self._trace_path(rv, path)
        plan = self._plan(rv.coordSys, reverse, start)
        if plan.struct.nstep == 0:
            return rv
        rays = rv._rays
        _lib.call("bt_trace_plan", plan.struct, rays, rays, _batoid_niter, _device.stream())
        rv.coordSys = plan.spec.coordSys_out
        return rv

    def _trace_path(self, rv, path):
        """Trace along ``path`` in place (optic.py:1052-1080)."""
        from .trace import reflect
        for item, turn, direction in self._path_steps(path):
            if turn:
                reflect(item.surface, rv, item.coordSys, obscuration=item.obscuration)
            else:
                item.trace(rv, reverse=(direction == "reverse"))
        return rv

    def traceFull(self, rv, reverse=False, path=None):
        """Per-Interface history {name: {'name', 'in', 'out'}}
        (optic.py:1085-1182).  With ``path`` the items are traced in that
        order (a turn back reflects off the item's surface), keyed
        ``name_0``, ``name_1``, ... for repeated visits."""
        from collections import OrderedDict
        from .plan import collect_steps
        out = OrderedDict()
        if path is None:
            for iface, _, _ in collect_steps(self, reverse):
                rin = rv.copy()
                iface.trace(rv, reverse=reverse)
                out[iface.name] = {'name': iface.name, 'in': rin, 'out': rv.copy()}
            return out
        from .trace import reflect
        rv_in = rv
        for item, turn, direction in self._path_steps(path):
            rv_out = rv_in.copy()
            if turn:
                reflect(item.surface, rv_out, item.coordSys, obscuration=item.obscuration)
            else:
                item.trace(rv_out, reverse=(direction == "reverse"))
            key = item.name + '_0'
            j = 1
            while key in out:
                key = item.name + '_{}'.format(j)
                j += 1
            out[key] = {'name': item.name, 'in': rv_in.copy(), 'out': rv_out.copy()}
            rv_in = rv_out
        return out

    def _path_steps(self, path):
        """(item, turns back here, direction after it) along ``path``
        (optic.py:1052-1080): the direction flips where the next item comes
        earlier in the nominal order than the current one."""
        nominalOrder = {}
        for name in path:
            if name not in nominalOrder:
                nominalOrder[name] = len(nominalOrder)
        direction = "forward"
        steps = []
        for i, name in enumerate(path):
            if i == len(path) - 1:
                nextDirection = "reverse" if nominalOrder[name] == 0 else direction
            else:
                nextDirection = ("reverse" if nominalOrder[path[i + 1]] < nominalOrder[name]
                                 else "forward")
            turn = nextDirection != direction
            direction = nextDirection
            steps.append((self[name], turn, direction))
        return steps

    def traceSplit(self, rv, minFlux=1e-3, reverse=False, _verbose=False):
        """Recursively split rays at every Interface into forward and reverse
        parts (optic.py:1184-1282); rays that are vignetted or fall below
        ``minFlux`` are dropped after each item (on the GPU: an
        order-preserving compaction).  Returns (forward, reverse) lists of
        RayVectors, each with its ``path``, in the reference's order."""
        if _verbose:
            s = "reverse" if reverse else "forward"
            print("traceSplit {}       {:30s} flux = {:18.8f}   nphot = {:10d}".format(
                s, self.name, *_live_stats(rv, None)))
        bound = None if _verbose else self._flux_bound(rv, minFlux)
        if bound is None:
            return self._split_batch([(rv, reverse, False)], minFlux, _verbose)[0]
        rv._flux_bound = bound
        try:
            return self._split_batch([(rv, reverse, False)], minFlux, _deferred=True)[0]
        finally:
            rv.__dict__.pop('_flux_bound', None)

    def _interfaces(self):
        for item in self.items:
            if isinstance(item, CompoundOptic):
                yield from item._interfaces()
            else:
                yield item

    def _flux_bound(self, rv, minFlux):
        """The largest |flux| among the unvignetted rays of ``rv``, when
        traceSplit may run without reading back counts (one read-back here;
        None otherwise).

        Every coating is constant, so a ray set's fluxes are bounded by this
        bound times the |transmissivity| and |reflectivity| factors along
        its path (rounded products are monotonic, so the bound computed in
        the same order holds bit for bit; vignetted rays stay vignetted); a
        set whose bound falls below minFlux holds no ray the filter keeps
        and is dropped unread.  That ends the recursion when minFlux > 0,
        every reflectivity that turns rays back (RefractiveInterface,
        Detector) is below 1 in magnitude, the bound is finite, and no
        unvignetted ray at or above minFlux is already failed (a failed
        ray's flux is not scaled)."""
        if not minFlux > 0:
            return None
        for item in self._interfaces():
            if item.skip or isinstance(item, (Mirror, Baffle)):
                continue
            if not isinstance(item, (RefractiveInterface, Detector)):
                return None
            for reverse in ((False,) if isinstance(item, Detector) else (False, True)):
                try:
                    ft, fr = (abs(float(f)) for f in item._split_fractions(reverse))
                except (AttributeError, TypeError, ValueError):
                    return None
                if not (fr < 1 and ft < float('inf')):
                    return None
        if len(rv) == 0:
            return 0.0
        flux = rv.torch("flux")
        unvig = rv.torch("vignetted") == 0
        top = torch.where(unvig & ~torch.isnan(flux), flux.abs(), torch.zeros_like(flux)).max()
        bad = (unvig & (flux >= minFlux) & (rv.torch("failed") != 0)).any()
        top, bad = torch.stack([top, bad.to(flux.dtype)]).tolist()
        if bad or not top < float('inf'):
            return None
        return top

    def _split_batch(self, inputs, minFlux, _verbose=False, _compact=True, _deferred=False):
        """traceSplit of several (rays, reverse, filtered) inputs at once.

        The reference pops one work item at a time from a LIFO stack and
        reads back the counts of its outputs before the next.  Here every
        pending item is processed in one round (items at the same index
        together, so a nested CompoundOptic also takes them in one call),
        with one read-back of the counts of all their outputs; each item
        carries its place in the reference's depth-first order -- the
        parent's key plus -p for its p-th pushed child (the last pushed is
        popped first) -- and the outputs, appended while their parent is
        processed, are sorted by (parent key, position) at the end, which
        reproduces the reference's lists exactly.

        The filter itself (unvignetted, flux >= minFlux) is only counted
        for the ray sets that go on to another item (by the split kernel
        itself where it can): they keep the dropped rays, which the next
        items skip (``_filtered``).  Only the sets that leave the outermost
        optic are copied out, kept rays in order (``_compact``; a nested
        optic returns its sets uncopied, with their counts), so each
        returned set holds exactly the rays the reference's per-item copies
        would.

        ``_deferred`` (the inputs carry ``_flux_bound``, see _flux_bound): no
        counts are read back between rounds; a set is dropped when its flux
        bound is below minFlux and otherwise traced on (if all its rays were
        vignetted, as a pass that skips every ray), and the sets leaving the
        outermost optic are counted once, at the end, the empty ones
        dropped."""
        if self.skip:
            for rv, _, _ in inputs:
                rv._live_slot = rv._live_count = None
            return [((None, rv) if rev else (rv, None)) for rv, rev, _ in inputs]
        if _deferred:
            return self._split_deferred(inputs, minFlux, _compact)
        last = len(self.items) - 1
        # work entries: (input, key, rays, reverse, item index, filtered)
        work = [(k, (), rv, rev, last if rev else 0, filt)
                for k, (rv, rev, filt) in enumerate(inputs)]
        outs = [([], []) for _ in inputs]              # per input: [(order, rays)] x 2
        while work:
            groups = {}
            for w in work:
                groups.setdefault(w[4], []).append(w)
            work = []
            children = []                              # (entry, side, rays, leaves)
            for index in sorted(groups):
                entries = groups[index]
                results = self.items[index]._split_batch(
                    [(w[2], w[3], w[5]) for w in entries], minFlux, _verbose, _compact=False)
                for w, (rF, rR) in zip(entries, results):
                    for side, rl in enumerate((rF, rR)):
                        if not isinstance(rl, list):   # a skipped item passes rays on
                            rl = [rl]
                        leaves = (side == 0 and w[4] == last) or (side == 1 and w[4] == 0)
                        children += [(w, side, rr, leaves) for rr in rl if rr]
            counts, copy = _count_children(children, minFlux, _compact)
            pushed, seq = {}, {}
            for ci, ((w, side, rr, leaves), m) in enumerate(zip(children, counts)):
                if m == 0:
                    continue
                k, key, _, _, index, _ = w
                parent = id(w)
                if leaves:
                    if _compact:
                        out = copy(ci, m)
                        out.path = rr.path
                    else:
                        out = rr
                        out._live_slot, out._live_count = None, m
                    s = seq.get(parent, 0)
                    seq[parent] = s + 1
                    outs[k][side].append(((key, s), out))
                else:
                    p = pushed.get(parent, 0)
                    pushed[parent] = p + 1
                    work.append((k, key + (-p,), rr, side == 1,
                                 index + 1 if side == 0 else index - 1, True))
        return [tuple([r for _, r in sorted(lst, key=lambda e: e[0])] for lst in o)
                for o in outs]

    def _split_deferred(self, inputs, minFlux, compact):
        """_split_batch without read-backs between rounds: the same rounds,
        keys and order, with a set dropped by its flux bound instead of its
        count (a set the bound keeps but the filter empties goes on empty,
        and leaves no output)."""
        last = len(self.items) - 1
        work = [(k, (), rv, rev, last if rev else 0, filt)
                for k, (rv, rev, filt) in enumerate(inputs)]
        outs = [([], []) for _ in inputs]
        leaving = []                                   # (input, side, order, rays)
        while work:
            groups = {}
            for w in work:
                groups.setdefault(w[4], []).append(w)
            work = []
            pushed, seq = {}, {}
            for index in sorted(groups):
                entries = groups[index]
                results = self.items[index]._split_batch(
                    [(w[2], w[3], w[5]) for w in entries], minFlux, _compact=False,
                    _deferred=True)
                for w, (rF, rR) in zip(entries, results):
                    k, key = w[0], w[1]
                    parent = id(w)
                    for side, rl in enumerate((rF, rR)):
                        if not isinstance(rl, list):   # a skipped item passes rays on
                            rl = [rl]
                        leaves = (side == 0 and w[4] == last) or (side == 1 and w[4] == 0)
                        for rr in rl:
                            if not rr or not rr._flux_bound >= minFlux:
                                continue
                            if leaves:
                                s = seq.get(parent, 0)
                                seq[parent] = s + 1
                                leaving.append((k, side, (key, s), rr))
                            else:
                                p = pushed.get(parent, 0)
                                pushed[parent] = p + 1
                                work.append((k, key + (-p,), rr, side == 1,
                                             w[4] + 1 if side == 0 else w[4] - 1, True))
        if compact and leaving:
            sel = RayVector._count_select([rr for _, _, _, rr in leaving], minFlux)
            for i, ((k, side, order, rr), m) in enumerate(zip(leaving, sel.counts.tolist())):
                if m:
                    out = sel.scatter(i, m)
                    out.path = rr.path
                    outs[k][side].append((order, out))
        else:
            for k, side, order, rr in leaving:
                rr._live_slot = rr._live_count = None
                outs[k][side].append((order, rr))
        return [tuple([r for _, r in sorted(lst, key=lambda e: e[0])] for lst in o)
                for o in outs]

    def clearObscuration(self, unless=()):
        for item in self.items:
            item.clearObscuration(unless=unless)

    def withSurface(self, name, surface):
        """Copy with the surface of item ``name`` replaced (optic.py:1695-1735)."""
        if name == self.name:
            raise ValueError("Cannot replace surface of CompoundOptic")
        item = self[name]
        return self._replaced(item, item.withSurface(surface))

    def _replaced(self, old, new):
        items = []
        for it in self.items:
            if it is old:
                items.append(new)
            elif isinstance(it, CompoundOptic):
                items.append(it._replaced(old, new))
            else:
                items.append(it)
        ret = self.__class__.__new__(self.__class__)
        ret.__init__(items, **_init_kwargs(self, ('items',)))
        return ret

    def withPerturbedSurface(self, name, perturbation):
        return self.withSurface(name, self[name].surface + perturbation)

    def withLocalRotation(self, rot, rotCenter=None, coordSys=None):
        """Copy rotated about this optic's own axes: the rotation
        R rot R^T in global axes, applied to every sub-item."""
        center = [0, 0, 0] if rotCenter is None else rotCenter
        cs = self.coordSys if coordSys is None else coordSys
        R = self.coordSys.rot
        return self.withGlobalRotation(R @ rot @ R.T, center, cs)

    def _full_name(self, name):
        name = self._names.get(name, name)
        if name not in self.itemDict:
            raise ValueError("Optic {} not found".format(name))
        return name

    def _with_item(self, name, change):
        """Copy with item ``name`` replaced by change(item) (the whole optic
        when ``name`` is this optic)."""
        name = self._full_name(name)
        if name == self.name:
            return change(self)
        item = self.itemDict[name]
        return self._replaced(item, change(item))

    def withGloballyShiftedOptic(self, name, shift):
        """Copy with item ``name`` shifted along global axes
        (optic.py:1410-1457)."""
        return self._with_item(name, lambda o: o.withGlobalShift(shift))

    def withLocallyShiftedOptic(self, name, shift):
        """Copy with item ``name`` shifted along its own axes
        (optic.py:1459-1481)."""
        item = self.itemDict[self._full_name(name)]
        return self.withGloballyShiftedOptic(name, item.coordSys.rot @ np.asarray(shift, float))

    def withGloballyRotatedOptic(self, name, rot, rotCenter=None, coordSys=None):
        """Copy with item ``name`` rotated about global axes; ``rotCenter``
        in ``coordSys`` (default global) (optic.py:1630-1693)."""
        return self._with_item(
            name, lambda o: o.withGlobalRotation(rot, rotCenter,
                                                 globalCoordSys if coordSys is None
                                                 else coordSys))

    def withLocallyRotatedOptic(self, name, rot, rotCenter=None, coordSys=None):
        """Copy with item ``name`` rotated about its own axes; ``rotCenter``
        in ``coordSys`` (default the item's frame) (optic.py:1567-1628)."""
        item = self.itemDict[self._full_name(name)]
        cs = item.coordSys if coordSys is None else coordSys
        return self._with_item(name, lambda o: o.withLocalRotation(rot, rotCenter, cs))

    def withInsertedOptic(self, before, item):
        """Copy with ``item`` inserted before the direct sub-item ``before``
        (optic.py:1760-1790)."""
        target = self._full_name(before)
        items = []
        for it in self.items:
            if self._names.get(it.name) == target:
                items.append(item)
            items.append(it)
        ret = self.__class__.__new__(self.__class__)
        ret.__init__(items, **_init_kwargs(self, ('items',)))
        return ret

    def withRemovedOptic(self, item):
        """Copy without the direct sub-item ``item`` (optic.py:1792-1818)."""
        target = self._full_name(item)
        items = [it for it in self.items if self._names.get(it.name) != target]
        ret = self.__class__.__new__(self.__class__)
        ret.__init__(items, **_init_kwargs(self, ('items',)))
        return ret

    def __eq__(self, other):
        if not self.__class__ == other.__class__:
            return False
        return (self.items == other.items and self.name == other.name
                and self.inMedium == other.inMedium and self.outMedium == other.outMedium
                and self.coordSys == other.coordSys)

    def __ne__(self, other):
        return not (self == other)

    def __hash__(self):
        return hash((self.__class__.__name__, self.items, self.name, self.inMedium,
                     self.outMedium, self.coordSys))

    def __repr__(self):
        out = "{!s}([".format(self.__class__.__name__)
        out += ",".join([repr(item) for item in self.items])
        return out + "]" + Optic._repr_helper(self) + ")"


class Lens(CompoundOptic):
    """Two Interfaces around a medium (optic.py:1821-1834)."""

    def __init__(self, items, **kwargs):
        CompoundOptic.__init__(self, items, **kwargs)
        assert len(items) == 2
        assert items[0].outMedium == items[1].inMedium
