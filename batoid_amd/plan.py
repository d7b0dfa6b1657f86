"""Optic -> trace plan compiler.

``CompoundOptic.trace`` (optic.py:1013-1049) visits its Interfaces in order
and each one transforms the rays into its frame, intersects/reflects/
refracts and then obscures (optic.py:391-424).  A plan is that sequence
flattened into tables the fused GPU kernel walks per ray:

* steps: action, surface index, outgoing-medium index, obscuration
  program, and the CoordTransform (dr, drot) from the previous step's frame;
* surfaces / media / obscuration records, shared between steps that use
  the same object (so a Bicubic table is uploaded once).
"""
import numpy as np

from . import _lib
from .coordTransform import CoordTransform


def _key(obj):
    """Table key: equal objects share one uploaded record set."""
    try:
        hash(obj)
        return (type(obj), obj)
    except TypeError:
        return (type(obj), id(obj))


def _start_of(start, iface):
    """(bt_start, packed network or None) of one step.

    ``start`` is "reference" (or None), "conic", or a mapping from
    Interface name to a trained NewtonStartNet; in a mapping, Newton
    surfaces without a network use the conic start."""
    from .newton_start import _is_newton
    if start is None or start == "reference":
        return _lib.START_REFERENCE, None
    if not _is_newton(iface.surface):
        return _lib.START_REFERENCE, None
    if start == "conic":
        return _lib.START_CONIC, None
    if isinstance(start, dict):
        net = start.get(iface.name)
        if net is None:
            return _lib.START_CONIC, None
        return _lib.START_MLP, net.packed
    raise ValueError(f"unknown start {start!r}: use 'reference', 'conic' or "
                     "a {name: NewtonStartNet} mapping")


def start_key(start):
    """Hashable identity of a start argument (for plan caches)."""
    if start is None or start == "reference":
        return "reference"
    if isinstance(start, dict):
        # the packed buffer changes whenever a network is re-fit
        return tuple(sorted((k, v.packed.data_ptr()) for k, v in start.items()))
    return start


_APB = (_lib.ASPHERE, _lib.POLYNOMIAL, _lib.BICUBIC)


def _sums_are_apb(recs):
    """True when every record is an analytic kind or belongs to a Sum of
    exactly (Asphere, Polynomial, Bicubic) children, in that order: the
    layout the trace's kKindsSumAPB kernel unrolls."""
    i = 0
    while i < len(recs):
        r = recs[i]
        if r.kind == _lib.SUM:
            kids = tuple(c.kind for c in recs[i + 1:i + 1 + r.nchild])
            if kids != _APB:
                return False
            i += 1 + r.nchild
        elif r.kind in _lib.ANALYTIC_KINDS:
            i += 1
        else:
            return False
    return True


def _has_nested_sum(recs):
    """True when a Sum record has a Sum among its children."""
    i = 0
    while i < len(recs):
        r = recs[i]
        if r.kind == _lib.SUM:
            c = i + 1
            for _ in range(r.nchild):
                if recs[c].kind == _lib.SUM:
                    return True
                c += 1
            i = c
        else:
            i += 1
    return False


class PlanSpec:
    """Host-side plan: ctypes records plus the numpy tables they point to."""

    def __init__(self, steps, coordSys_in, start=None):
        self.surf_recs, self.surf_arrs = [], []
        self.med_recs, self.med_arrs = [], []
        self.ob_recs, self.ob_arrs = [], []
        self.steps = []
        self.interfaces = []
        surf_index, med_index, ob_index = {}, {}, {}
        refracts = False
        fast = mlp = False
        self.networks = []      # packed network buffers the steps point to
        cs = coordSys_in
        for iface, action, m2 in steps:
            key = _key(iface.surface)
            if key not in surf_index:
                surf_index[key] = len(self.surf_recs)
                r, a = iface.surface._records()
                self.surf_recs += r
                self.surf_arrs += a
            mi = -1
            if action == _lib.ACT_REFRACT:
                refracts = True
                mkey = _key(m2)
                if mkey not in med_index:
                    med_index[mkey] = len(self.med_recs)
                    r, a = m2._records()
                    self.med_recs += r
                    self.med_arrs += a
                mi = med_index[mkey]
            oi, ol = -1, 0
            if iface.obscuration is not None:
                okey = _key(iface.obscuration)
                if okey not in ob_index:
                    r, a = iface.obscuration._records()
                    ob_index[okey] = (len(self.ob_recs), len(r))
                    self.ob_recs += r
                    self.ob_arrs += a
                oi, ol = ob_index[okey]
            ct = CoordTransform(cs, iface.coordSys)
            st = _lib.BtStep()
            st.action = action
            st.surface = surf_index[key]
            st.medium = mi
            st.obsc = oi
            st.obsc_len = ol
            mode, packed = _start_of(start, iface)
            st.start = mode
            if packed is not None:
                self.networks.append(packed)
                st.mlp = packed.data_ptr()
            fast = fast or mode != _lib.START_REFERENCE
            mlp = mlp or mode == _lib.START_MLP
            for i, v in enumerate(ct.dr):
                st.dr[i] = float(v)
            for i, v in enumerate(np.asarray(ct.drot).ravel()):
                st.drot[i] = float(v)
            self.steps.append(st)
            self.interfaces.append(iface)
            cs = iface.coordSys
        self.coordSys_in = coordSys_in
        self.coordSys_out = cs
        analytic = all(r.kind in _lib.ANALYTIC_KINDS for r in self.surf_recs)
        self.flags = ((0 if refracts else _lib.PLAN_NO_REFRACT)
                      | (_lib.PLAN_FAST_START if fast else 0)
                      | (_lib.PLAN_MLP_START if mlp else 0)
                      | (_lib.PLAN_ANALYTIC if analytic else 0)
                      | (_lib.PLAN_SUM_APB if not analytic and _sums_are_apb(self.surf_recs)
                         else 0))
        if not self.flags & (_lib.PLAN_ANALYTIC | _lib.PLAN_SUM_APB) \
                and not _has_nested_sum(self.surf_recs):
            self.flags |= _lib.PLAN_FLAT_SUMS


class DevicePlan:
    """A PlanSpec uploaded to GPU memory; ``struct`` is the bt_plan."""

    def __init__(self, spec):
        from ._device import Materialized, upload_bytes
        self.spec = spec
        self.surfaces = Materialized(spec.surf_recs, spec.surf_arrs) if spec.surf_recs else None
        self.media = Materialized(spec.med_recs, spec.med_arrs) if spec.med_recs else None
        self.obscs = Materialized(spec.ob_recs, spec.ob_arrs) if spec.ob_recs else None
        self.steps = upload_bytes(_lib.struct_array_bytes(spec.steps))
        p = _lib.BtPlan()
        p.nstep = len(spec.steps)
        p.flags = spec.flags
        p.nmedia = len(spec.med_recs)
        p.steps = self.steps.data_ptr() if spec.steps else None
        p.surfaces = self.surfaces.buffer.data_ptr() if self.surfaces else None
        p.media = self.media.buffer.data_ptr() if self.media else None
        p.obscs = self.obscs.buffer.data_ptr() if self.obscs else None
        self.struct = p


def collect_steps(optic, reverse=False):
    """(interface, action, outgoing medium) for every unskipped Interface,
    in trace order (optic.py:1045-1049, 541-543, 776-781, 814-816)."""
    from .optic import CompoundOptic, Interface, Mirror, RefractiveInterface, OPDScreen
    out = []

    def visit(item):
        if isinstance(item, CompoundOptic):
            items = item.items if not reverse else tuple(reversed(item.items))
            for sub in items:
                if not sub.skip:
                    visit(sub)
        elif isinstance(item, OPDScreen):
            raise NotImplementedError("OPDScreen is outside the GPU trace path")
        elif isinstance(item, RefractiveInterface):
            m2 = item.inMedium if reverse else item.outMedium
            out.append((item, _lib.ACT_REFRACT, m2))
        elif isinstance(item, Mirror):
            out.append((item, _lib.ACT_REFLECT, None))
        elif isinstance(item, Interface):
            out.append((item, _lib.ACT_INTERSECT, None))
        else:
            raise TypeError(f"cannot trace {type(item).__name__}")

    visit(optic)
    return out
