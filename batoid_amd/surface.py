"""Surfaces (batoid/surface.py; device math follows src/<surface>.cpp).

Each surface describes itself as ``bt_surface`` records whose derived
constants are computed here exactly as the reference constructors compute
them (same operations, same order, IEEE double), so the device evaluates
bit-identical sags, normals and intersections."""
from numbers import Integral, Real

import numpy as np

from . import _lib
from ._cache import DeviceCached
from .zernike_poly import zernike_coef_arrays


def _f(x):
    return np.float64(x)


def _div(a, b):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.float64(a) / np.float64(b)


def _rec(kind, p=(), n1=0, n2=0, flags=0, nchild=0):
    r = _lib.BtSurface()
    r.kind = kind
    r.nchild = nchild
    r.n1 = n1
    r.n2 = n2
    r.flags = flags
    for i, v in enumerate(p):
        r.p[i] = float(v)
    return r


def _quadric_params(R, conic):
    # quadric.cpp:9-16, evaluated in the same order
    R = _f(R)
    cp1 = _f(conic) + 1
    return [R, _f(conic), R * R, _div(_div(1., R), R), cp1, _div(1., cp1),
            _div(R, cp1), _div(_div(R * R, cp1), cp1), _div(_div(cp1, R), R)]


class Surface(DeviceCached):
    """A 2-D surface z(x, y) with sag, normal and ray interactions."""

    def _leaf_record(self):
        raise NotImplementedError

    def _records(self):
        rec, arrs = self._leaf_record()
        return [rec], [arrs]

    def _eval_args(self, x, y):
        from . import _device
        xb, yb = np.broadcast_arrays(np.atleast_1d(np.asarray(x, float)),
                                     np.atleast_1d(np.asarray(y, float)))
        return xb.shape, _device.f64(xb.ravel()), _device.f64(yb.ravel())

    def sag(self, x, y):
        """Surface height z(x, y), evaluated on the GPU."""
        from . import _device
        scalar = np.ndim(x) == 0 and np.ndim(y) == 0
        shape, xd, yd = self._eval_args(x, y)
        out = _device.empty_f64(xd.numel())
        _lib.call("bt_surface_sag", self._dev.ptr(), _device.ptr(xd), _device.ptr(yd),
                  _device.ptr(out), xd.numel(), _device.stream())
        res = _device.to_numpy(out)
        return float(res[0]) if scalar else res.reshape(shape)

    def normal(self, x, y):
        """Unit normal at (x, y, z(x, y)); shape (..., 3), or (3,) for
        scalar inputs."""
        from . import _device
        scalar = np.ndim(x) == 0 and np.ndim(y) == 0
        shape, xd, yd = self._eval_args(x, y)
        n = xd.numel()
        out = _device.empty_f64(3 * n)
        _lib.call("bt_surface_normal", self._dev.ptr(), _device.ptr(xd),
                  _device.ptr(yd), _device.ptr(out), n, _device.stream())
        res = _device.to_numpy(out).reshape(3, n).T
        return res[0] if scalar else res.reshape(shape + (3,))

    def intersect(self, rv, coordSys=None, coating=None):
        """Propagate rays to this surface in place (trace.py:47-81)."""
        from .trace import intersect
        return intersect(self, rv, coordSys, coating)

    def reflect(self, rv, coordSys=None, coating=None):
        """Intersect and reflect rays in place (trace.py:84-98)."""
        from .trace import reflect
        return reflect(self, rv, coordSys, coating)

    def rSplit(self, rv, inMedium, outMedium, coating, coordSys=None, obscuration=None,
               _live_min_flux=None, _live_counts=None, _no_split=False):
        """Refracted (``rv``, in place) and reflected (new) rays at this
        surface, fluxes split by ``coating`` (surface.py:107-133)."""
        from .trace import rSplit
        return rSplit(self, rv, inMedium, outMedium, coating, coordSys, obscuration,
                      _live_min_flux, _live_counts, _no_split)

    @property
    def _analytic(self):
        """Every record of this surface is a Plane, Sphere, Paraboloid,
        Quadric, Asphere or Tilted."""
        a = self.__dict__.get("_analytic_c")
        if a is None:
            a = all(r.kind in _lib.ANALYTIC_KINDS for r in self._records()[0])
            self.__dict__["_analytic_c"] = a
        return a

    def refract(self, rv, inMedium, outMedium, coordSys=None, coating=None):
        """Intersect and refract rays in place (trace.py:101-116)."""
        from .trace import refract
        return refract(self, rv, inMedium, outMedium, coordSys, coating)

    def __add__(self, rhs):
        return Sum(self, rhs)


class Plane(Surface):
    """z = 0 (plane.cpp)."""

    def _leaf_record(self):
        return _rec(_lib.PLANE), (None,) * 4

    def __hash__(self):
        return hash("batoid.Plane")

    def __eq__(self, rhs):
        return isinstance(rhs, Plane)

    def __repr__(self):
        return "Plane()"


class Tilted(Surface):
    """z = x tanx + y tany (tilted.cpp)."""

    def __init__(self, tanx, tany):
        self.tanx = tanx
        self.tany = tany

    def _leaf_record(self):
        return _rec(_lib.TILTED, [self.tanx, self.tany]), (None,) * 4

    def __hash__(self):
        return hash(("batoid.Tilted", self.tanx, self.tany))

    def __eq__(self, rhs):
        return isinstance(rhs, Tilted) and self.tanx == rhs.tanx and self.tany == rhs.tany

    def __repr__(self):
        return f"Tilted({self.tanx}, {self.tany})"


class Paraboloid(Surface):
    """z = r^2 / (2R) (paraboloid.cpp)."""

    def __init__(self, R):
        self.R = R

    def _leaf_record(self):
        # paraboloid.cpp:9-11: _Rinv(1./R), _2Rinv(1./2/R)
        return (_rec(_lib.PARABOLOID, [self.R, _div(1., self.R), _div(_div(1., 2), self.R)]),
                (None,) * 4)

    def __hash__(self):
        return hash(("batoid.Paraboloid", self.R))

    def __eq__(self, rhs):
        return isinstance(rhs, Paraboloid) and self.R == rhs.R

    def __repr__(self):
        return f"Paraboloid({self.R})"


class Sphere(Surface):
    """z = R (1 - sqrt(1 - r^2/R^2)) (sphere.cpp)."""

    def __init__(self, R):
        self.R = R

    def _leaf_record(self):
        R = _f(self.R)
        # sphere.cpp:9-11: _Rsq(R*R), _Rinv(1./R), _Rinvsq(1./R/R)
        return (_rec(_lib.SPHERE, [R, R * R, _div(1., R), _div(_div(1., R), R)]),
                (None,) * 4)

    def __hash__(self):
        return hash(("batoid.Sphere", self.R))

    def __eq__(self, rhs):
        return isinstance(rhs, Sphere) and self.R == rhs.R

    def __repr__(self):
        return f"Sphere({self.R})"


class Quadric(Surface):
    """Conic of revolution with vertex radius R and conic constant k
    (quadric.cpp)."""

    def __init__(self, R, conic):
        self.R = R
        self.conic = conic

    def _leaf_record(self):
        return _rec(_lib.QUADRIC, _quadric_params(self.R, self.conic)), (None,) * 4

    def __hash__(self):
        return hash(("batoid.Quadric", self.R, self.conic))

    def __eq__(self, rhs):
        return isinstance(rhs, Quadric) and self.R == rhs.R and self.conic == rhs.conic

    def __repr__(self):
        return f"Quadric({self.R}, {self.conic})"


class Asphere(Surface):
    """Quadric plus even polynomial sum_i alpha_i r^(2i), with coefs[0] the
    coefficient of r^(2 imin) (surface.py:342-430, asphere.cpp)."""

    def __init__(self, R, conic, coefs, imin=2):
        if not isinstance(imin, Integral):
            raise TypeError("imin must be an int")
        if imin <= 0:
            raise ValueError("imin must be >= 1")
        if not all(isinstance(c, Real) for c in coefs):
            raise TypeError("coefs must be a list of floats")
        self.R = R
        self.conic = conic
        self.coefs = coefs
        self.imin = imin
        # index 0 of the padded array is the r^2 term (surface.py:391-399)
        padded = np.concatenate([np.zeros(imin - 1), coefs]) if imin != 1 else coefs
        self._padded_coefs = np.ascontiguousarray(padded, dtype=float)

    def _leaf_record(self):
        c = self._padded_coefs
        # asphere.cpp:10-15: dz/dr coefficients coefs[j] * (2, 4, 6, ...)
        dzdr = np.array([c[j] * float(2 * (j + 1)) for j in range(len(c))], dtype=float)
        return (_rec(_lib.ASPHERE, _quadric_params(self.R, self.conic), n1=len(c)),
                (c, dzdr, None, None))

    def __hash__(self):
        return hash(("batoid.Asphere", self.R, self.conic, tuple(self.coefs), self.imin))

    def __eq__(self, rhs):
        return (isinstance(rhs, Asphere) and self.R == rhs.R and self.conic == rhs.conic
                and np.array_equal(self.coefs, rhs.coefs) and self.imin == rhs.imin)

    def __repr__(self):
        out = f"Asphere({self.R}, {self.conic}, {self.coefs!r}"
        return out + (")" if self.imin == 2 else f", imin={self.imin})")


def _horner_trim(tables):
    """Trim table of 2-D Horner coefficient tables, as the device reads each
    one (row i = flat[i*nx:(i+1)*nx], ny rows): its last row holding a
    coefficient other than +0.0 (-1 if none), then every row's length up to
    its last such coefficient; the tables' entries concatenated (int32).
    Only +0.0 counts as trimmable: -0.0 is kept, since the loop's result can
    depend on it."""
    out = []
    for flat, nx, ny in tables:
        flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
        if ny > 0 and nx > 0:
            rows = flat[:nx * ny].reshape(ny, nx).view(np.uint64) != 0
            lens = [int(np.flatnonzero(row)[-1]) + 1 if row.any() else 0 for row in rows]
        else:
            lens = [0] * max(ny, 0)
        last = max([i for i, n in enumerate(lens) if n > 0], default=-1)
        out += [last] + lens
    return np.array(out, dtype=np.int32)


class Zernike(Surface):
    """Annular Zernike series sum_j coef[j] Z_j(x/R_outer, y/R_outer) in
    Noll order (surface.py:433-518), evaluated as a 2-D polynomial
    (polynomialSurface.cpp)."""

    def __init__(self, coef, R_outer=1.0, R_inner=0.0, x_origin=0.0, y_origin=0.0):
        self.coef = np.array(coef, dtype=float, order="C")
        self.R_outer = float(R_outer)
        self.R_inner = float(R_inner)
        self.x_origin = float(x_origin)
        self.y_origin = float(y_origin)
        c, gx, gy = zernike_coef_arrays(self.coef, self.R_outer, self.R_inner)
        self._xycoef = np.ascontiguousarray(c)
        self._xycoef_gradx = np.ascontiguousarray(gx)
        self._xycoef_grady = np.ascontiguousarray(gy)

    def _leaf_record(self):
        c = self._xycoef
        n1, n2 = c.shape
        trim = _horner_trim([(c, n1, n2), (self._xycoef_gradx, n1 - 1, n2 - 1),
                             (self._xycoef_grady, n1 - 1, n2 - 1)])
        return (_rec(_lib.POLYNOMIAL, [self.x_origin, self.y_origin], n1=n1, n2=n2),
                (c, self._xycoef_gradx, self._xycoef_grady, trim))

    def __hash__(self):
        return hash(("batoid.Zernike", tuple(self.coef), self.R_outer, self.R_inner,
                     self.x_origin, self.y_origin))

    def __eq__(self, rhs):
        return (isinstance(rhs, Zernike) and np.array_equal(self.coef, rhs.coef)
                and self.R_outer == rhs.R_outer and self.R_inner == rhs.R_inner
                and self.x_origin == rhs.x_origin and self.y_origin == rhs.y_origin)

    def __repr__(self):
        out = f"Zernike({self.coef!r}"
        if self.R_outer != 1.0:
            out += f", R_outer={self.R_outer}"
        if self.R_inner != 0.0:
            out += f", R_inner={self.R_inner}"
        if self.x_origin != 0.0:
            out += f", x_origin={self.x_origin}"
        if self.y_origin != 0.0:
            out += f", y_origin={self.y_origin}"
        return out + ")"


class Bicubic(Surface):
    """Bicubic (Hermite) interpolation of a regular grid (surface.py:521-663,
    table.cpp).  Missing derivative tables are built by central differences
    (one-sided at the edges) exactly as the reference wrapper does."""

    def __init__(self, xs, ys, zs, dzdxs=None, dzdys=None, d2zdxdys=None,
                 nanpolicy='nan'):
        assert nanpolicy.upper() in ['NAN', 'ZERO']
        self._xs = np.array(xs, dtype=float, order="C")
        self._ys = np.array(ys, dtype=float, order="C")
        self._zs = np.array(zs, dtype=float, order="C")
        self._x0 = xs[0]
        self._y0 = ys[0]
        dx = self._dx = (self._xs[-1] - self._xs[0]) / (len(self._xs) - 1)
        dy = self._dy = (self._ys[-1] - self._ys[0]) / (len(self._ys) - 1)
        z = self._zs
        if dzdxs is None:
            dzdxs = np.empty_like(z)
            dzdxs[:, 1:-1] = (z[:, 2:] - z[:, :-2]) / (2 * dx)
            dzdxs[:, 0] = (z[:, 1] - z[:, 0]) / dx
            dzdxs[:, -1] = (z[:, -1] - z[:, -2]) / dx
        if dzdys is None:
            dzdys = np.empty_like(z)
            dzdys[1:-1, :] = (z[2:, :] - z[:-2, :]) / (2 * dy)
            dzdys[0, :] = (z[1, :] - z[0, :]) / dy
            dzdys[-1, :] = (z[-1, :] - z[-2, :]) / dy
        if d2zdxdys is None:
            d2zdxdys = np.empty_like(z)
            d2zdxdys[:, 1:-1] = (dzdys[:, 2:] - dzdys[:, :-2]) / (2 * dx)
            d2zdxdys[:, 0] = (dzdys[:, 1] - dzdys[:, 0]) / dx
            d2zdxdys[:, -1] = (dzdys[:, -1] - dzdys[:, -2]) / dx
        self._dzdxs = np.array(dzdxs, dtype=float, order="C")
        self._dzdys = np.array(dzdys, dtype=float, order="C")
        self._d2zdxdys = np.array(d2zdxdys, dtype=float, order="C")
        self.nanpolicy = nanpolicy

    xs = property(lambda self: self._xs)
    ys = property(lambda self: self._ys)
    zs = property(lambda self: self._zs)
    dzdxs = property(lambda self: self._dzdxs)
    dzdys = property(lambda self: self._dzdys)
    d2zdxdys = property(lambda self: self._d2zdxdys)

    def _leaf_record(self):
        use_nan = 1 if self.nanpolicy.upper() == 'NAN' else 0
        return (_rec(_lib.BICUBIC, [self._x0, self._y0, self._dx, self._dy],
                     n1=len(self._xs), n2=len(self._ys), flags=use_nan),
                (self._zs, self._dzdxs, self._dzdys, self._d2zdxdys))

    def __hash__(self):
        return hash(("Bicubic", tuple(self.xs), tuple(self.ys), tuple(self.zs.ravel()),
                     self.nanpolicy))

    def __eq__(self, rhs):
        return (isinstance(rhs, Bicubic) and np.array_equal(self.xs, rhs.xs)
                and np.array_equal(self.ys, rhs.ys) and np.array_equal(self.zs, rhs.zs)
                and np.array_equal(self.dzdxs, rhs.dzdxs)
                and np.array_equal(self.dzdys, rhs.dzdys)
                and np.array_equal(self.d2zdxdys, rhs.d2zdxdys)
                and self.nanpolicy.upper() == rhs.nanpolicy.upper())

    def __repr__(self):
        out = f"Bicubic({self.xs!r}, {self.ys!r}, {self.zs!r}, "
        out += f"{self.dzdxs!r}, {self.dzdys!r}, {self.d2zdxdys!r}"
        return out + (")" if self.nanpolicy.upper() == "NAN" else ", nanpolicy='zero')")


class Sum(Surface):
    """Sum of surfaces; intersection seeds Newton with the first surface's
    intersection (surface.py:666-710, sum.cpp).  Any number of children;
    children may themselves be Sums, nested at most MAX_SUM_DEPTH levels
    below this one (the GPU evaluates nested Sums to that depth)."""

    def __init__(self, *args):
        from collections.abc import Sequence
        if len(args) == 1 and isinstance(args[0], Sequence):
            args = args[0]
        assert all(isinstance(arg, Surface) for arg in args)
        self.surfaces = tuple(args)

    def _depth(self):
        return 1 + max((s._depth() for s in self.surfaces if isinstance(s, Sum)), default=0)

    def _records(self):
        """The Sum record, then each child's records in order; n1 counts the
        records below this one, so a child Sum's subtree can be stepped
        over."""
        if self._depth() > 1 + _lib.MAX_SUM_DEPTH:
            raise NotImplementedError(
                f"Sums nested more than {_lib.MAX_SUM_DEPTH} levels deep")
        recs, arrs = [], []
        for s in self.surfaces:
            r, a = s._records()
            recs += r
            arrs += a
        head = _rec(_lib.SUM, nchild=len(self.surfaces), n1=len(recs))
        return [head] + recs, [(None,) * 4] + arrs

    def __hash__(self):
        return hash(("batoid.Sum", tuple(self.surfaces)))

    def __eq__(self, rhs):
        return isinstance(rhs, Sum) and self.surfaces == rhs.surfaces

    def __repr__(self):
        return f"Sum({self.surfaces})"
