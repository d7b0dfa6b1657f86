"""Obscurations (batoid/obscuration.py; kernels follow src/obscuration.cpp).

An obscuration compiles to a postfix program of ``bt_obsc`` records that the
trace kernels evaluate per ray; composite obscurations (Negation, Union,
Intersection) append their operator after their operands' programs."""
import math

import numpy as np

from . import _lib
from ._cache import DeviceCached


def _op(op, p=(), n=0):
    r = _lib.BtObsc()
    r.op = op
    r.n = n
    for i, v in enumerate(p):
        r.p[i] = float(v)
    return r


class Obscuration(DeviceCached):
    """Marks rays as vignetted where ``contains(x, y)`` is true (local x, y;
    z is ignored)."""

    def _program(self):
        """(records, arrays) of this obscuration's postfix program."""
        raise NotImplementedError

    def _records(self):
        recs, arrs = self._program()
        depth = _stack_depth(recs)
        if depth > _lib.OBSC_MAX_STACK:
            raise ValueError(f"obscuration nests too deeply ({depth} > {_lib.OBSC_MAX_STACK})")
        return recs, arrs

    def contains(self, x, y):
        """Whether (x, y) is obscured, evaluated on the GPU."""
        from . import _device
        scalar = np.ndim(x) == 0 and np.ndim(y) == 0
        xb, yb = np.broadcast_arrays(np.atleast_1d(x), np.atleast_1d(y))
        xd, yd = _device.f64(xb.ravel()), _device.f64(yb.ravel())
        out = _device.u8(np.zeros(xd.numel(), dtype=bool))
        dev = self._dev
        _lib.call("bt_obsc_contains", dev.ptr(), dev.count, _device.ptr(xd),
                  _device.ptr(yd), _device.ptr(out), xd.numel(), _device.stream())
        res = _device.to_numpy(out).astype(bool)
        return bool(res[0]) if scalar else res.reshape(xb.shape)

    def obscure(self, rv):
        """Vignette, in place, the rays of ``rv`` that this obscures."""
        from .trace import obscure
        return obscure(self, rv)


def _stack_depth(recs):
    sp = depth = 0
    for r in recs:
        if r.op <= _lib.OB_POLYGON:
            sp += 1
        elif r.op in (_lib.OB_UNION, _lib.OB_INTERSECTION):
            sp -= r.n - 1
        depth = max(depth, sp)
    return depth


class ObscCircle(Obscuration):
    """Circle of ``radius`` centred at (x, y) (obscuration.cpp:32-34)."""

    def __init__(self, radius, x=0.0, y=0.0):
        self.radius = radius
        self.x = x
        self.y = y

    def _program(self):
        return [_op(_lib.OB_CIRCLE, [self.radius, self.x, self.y])], [(None, None)]

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and self.radius == rhs.radius
                and self.x == rhs.x and self.y == rhs.y)

    def __hash__(self):
        return hash(("batoid.ObscCircle", self.radius, self.x, self.y))

    def __repr__(self):
        out = f"ObscCircle({self.radius}"
        if self.x != 0 or self.y != 0:
            out += f", {self.x}, {self.y}"
        return out + ")"


class ObscAnnulus(Obscuration):
    """Annulus inner <= r < outer centred at (x, y) (obscuration.cpp:80-83)."""

    def __init__(self, inner, outer, x=0.0, y=0.0):
        self.inner = inner
        self.outer = outer
        self.x = x
        self.y = y

    def _program(self):
        return ([_op(_lib.OB_ANNULUS, [self.inner, self.outer, self.x, self.y])],
                [(None, None)])

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and self.inner == rhs.inner
                and self.outer == rhs.outer and self.x == rhs.x and self.y == rhs.y)

    def __hash__(self):
        return hash(("batoid.ObscAnnulus", self.inner, self.outer, self.x, self.y))

    def __repr__(self):
        out = f"ObscAnnulus({self.inner}, {self.outer}"
        if self.x != 0 or self.y != 0:
            out += f", {self.x}, {self.y}"
        return out + ")"


class ObscRectangle(Obscuration):
    """Rectangle width x height centred at (x, y), rotated counter-clockwise
    by ``theta`` (obscuration.cpp:108-121)."""

    def __init__(self, width, height, x=0.0, y=0.0, theta=0.0):
        self.width = width
        self.height = height
        self.x = x
        self.y = y
        self.theta = theta

    def _program(self):
        # cos/sin from libm, as the reference constructor (obscuration.cpp:110-111)
        c, s = math.cos(self.theta), math.sin(self.theta)
        return ([_op(_lib.OB_RECTANGLE, [self.width, self.height, self.x, self.y,
                                         c, s, self.theta])], [(None, None)])

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and self.width == rhs.width
                and self.height == rhs.height and self.x == rhs.x
                and self.y == rhs.y and self.theta == rhs.theta)

    def __hash__(self):
        return hash(("batoid.ObscRectangle", self.width, self.height, self.x,
                     self.y, self.theta))

    def __repr__(self):
        out = f"ObscRectangle({self.width}, {self.height}"
        if self.x != 0.0 or self.y != 0.0:
            out += f", {self.x}, {self.y}"
        if self.theta != 0.0:
            out += f", theta={self.theta}"
        return out + ")"


class ObscRay(Obscuration):
    """Half-infinite strip of ``width`` from (x, y) in direction ``theta``
    (obscuration.cpp:146-159)."""

    def __init__(self, width, theta, x=0.0, y=0.0):
        self.width = width
        self.theta = theta
        self.x = x
        self.y = y

    def _program(self):
        c, s = math.cos(self.theta), math.sin(self.theta)
        return ([_op(_lib.OB_RAY, [self.width, self.x, self.y, c, s, self.theta])],
                [(None, None)])

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and self.width == rhs.width
                and self.theta == rhs.theta and self.x == rhs.x and self.y == rhs.y)

    def __hash__(self):
        return hash(("batoid.ObscRay", self.width, self.theta, self.x, self.y))

    def __repr__(self):
        out = f"ObscRay({self.width}, {self.theta}"
        if self.x != 0.0 or self.y != 0.0:
            out += f", {self.x}, {self.y}"
        return out + ")"


class ObscPolygon(Obscuration):
    """Polygon with vertices (xs[i], ys[i]), crossing-number rule
    (obscuration.cpp:201-225)."""

    def __init__(self, xs, ys):
        self.xs = np.ascontiguousarray(xs, dtype=float)
        self.ys = np.ascontiguousarray(ys, dtype=float)

    def _program(self):
        return [_op(_lib.OB_POLYGON, n=len(self.xs))], [(self.xs, self.ys)]

    def containsGrid(self, xgrid, ygrid):
        """contains() on the (ny, nx) grid of ``xgrid`` x ``ygrid``."""
        xx, yy = np.meshgrid(np.asarray(xgrid, float), np.asarray(ygrid, float))
        return self.contains(xx, yy)

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and np.all(self.xs == rhs.xs)
                and np.all(self.ys == rhs.ys))

    def __hash__(self):
        return hash(("batoid.ObscPolygon", tuple(self.xs), tuple(self.ys)))

    def __repr__(self):
        return f"ObscPolygon({self.xs!r}, {self.ys!r})"


class ObscNegation(Obscuration):
    """Obscures exactly what ``original`` leaves clear (obscuration.cpp:320-322)."""

    def __init__(self, original):
        self.original = original

    def _program(self):
        recs, arrs = self.original._program()
        return recs + [_op(_lib.OB_NEGATION)], arrs + [(None, None)]

    def __eq__(self, rhs):
        return type(rhs) == type(self) and self.original == rhs.original

    def __hash__(self):
        return hash(("batoid.ObscNegation", self.original))

    def __repr__(self):
        return f"ObscNegation({self.original!r})"


def _items(items):
    if len(items) == 0:
        raise ValueError("Not enough items")
    if len(items) == 1:
        if isinstance(items[0], Obscuration):
            items = [items[0]]
        elif isinstance(items[0], (list, tuple)):
            items = items[0]
        else:
            raise TypeError("Single input argument must be an Obscuration or a list of them.")
    if not all(isinstance(item, Obscuration) for item in items):
        raise TypeError("All items must be Obscurations")
    return list(items)


class _Composite(Obscuration):
    _OP = None

    def __init__(self, *items):
        items = _items(items)
        self._order = items              # evaluation order, as passed
        self.items = sorted(items, key=repr)

    def _program(self):
        recs, arrs = [], []
        for item in self._order:
            r, a = item._program()
            recs += r
            arrs += a
        return recs + [_op(self._OP, n=len(self._order))], arrs + [(None, None)]

    def __eq__(self, rhs):
        return type(rhs) == type(self) and self.items == rhs.items

    def __hash__(self):
        return hash((f"batoid.{type(self).__name__}", tuple(self.items)))

    def __repr__(self):
        return f"{type(self).__name__}(" + ", ".join(repr(i) for i in self.items) + ")"


class ObscUnion(_Composite):
    """Obscured where any item obscures (obscuration.cpp:364-370)."""
    _OP = _lib.OB_UNION


class ObscIntersection(_Composite):
    """Obscured where every item obscures (obscuration.cpp:415-421)."""
    _OP = _lib.OB_INTERSECTION
