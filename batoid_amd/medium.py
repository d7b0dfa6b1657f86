"""Refractive media (batoid/medium.py; kernels follow src/medium.cpp).

Each medium describes itself as one ``bt_medium`` record; ``getN`` evaluates
n(lambda) on the GPU."""
import numpy as np

from . import _lib
from ._cache import DeviceCached


class Medium(DeviceCached):
    """A refractive medium: ``getN(wavelength)`` with wavelength in metres."""

    def _record(self):
        raise NotImplementedError

    def _records(self):
        rec, arrs = self._record()
        return [rec], [arrs]

    def getN(self, wavelength):
        """Refractive index at vacuum ``wavelength`` (metres), on the GPU."""
        from . import _device
        scalar = np.ndim(wavelength) == 0
        w = _device.f64(np.atleast_1d(wavelength) if not hasattr(wavelength, "device") else wavelength)
        out = _device.empty_f64(w.numel())
        _lib.call("bt_medium_get_n", self._dev.ptr(), _device.ptr(w), _device.ptr(out),
                  w.numel(), _device.stream())
        res = _device.to_numpy(out)
        return float(res[0]) if scalar else res.reshape(np.shape(wavelength))


def _rec(kind, p=(), n=0):
    r = _lib.BtMedium()
    r.kind = kind
    r.n = n
    for i, v in enumerate(p):
        r.p[i] = float(v)
    return r


class ConstMedium(Medium):
    """Wavelength-independent index ``n`` (medium.cpp:30-32)."""

    def __init__(self, n):
        self.n = n

    def _record(self):
        return _rec(_lib.MED_CONST, [self.n]), (None, None)

    def __eq__(self, rhs):
        return type(rhs) == type(self) and self.n == rhs.n

    def __hash__(self):
        return hash(("batoid.ConstMedium", self.n))

    def __repr__(self):
        return f"ConstMedium({self.n})"


class TableMedium(Medium):
    """Linear interpolation in a table of (wavelength [m], n); NaN outside
    the table (medium.cpp:89-105)."""

    def __init__(self, wavelengths, ns):
        self.wavelengths = np.array(wavelengths, dtype=float)
        self.ns = np.array(ns, dtype=float)

    @classmethod
    def fromTxt(cls, filename, **kwargs):
        """Two-column text file: wavelength in microns, index."""
        wavelength, n = np.loadtxt(filename, unpack=True, **kwargs)
        return TableMedium(wavelength * 1e-6, n)

    def _record(self):
        return (_rec(_lib.MED_TABLE, n=len(self.wavelengths)),
                (self.wavelengths, self.ns))

    def __eq__(self, rhs):
        return (type(rhs) == type(self)
                and np.array_equal(self.wavelengths, rhs.wavelengths)
                and np.array_equal(self.ns, rhs.ns))

    def __hash__(self):
        return hash(("batoid.TableMedium", tuple(self.wavelengths), tuple(self.ns)))

    def __repr__(self):
        return f"TableMedium({self.wavelengths!r}, {self.ns!r})"


def _six(args, kwargs, names):
    if len(args) == 6:
        return tuple(args)
    if len(args) == 1:
        return tuple(args[0])
    if kwargs:
        return tuple(kwargs[k] for k in names)
    raise ValueError("Incorrect number of arguments")


class SellmeierMedium(Medium):
    """Sellmeier dispersion, coefficients (B1, B2, B3, C1, C2, C3) with the
    wavelength in microns (medium.cpp:145-149)."""

    def __init__(self, *args, **kwargs):
        self.coefs = _six(args, kwargs, ['B1', 'B2', 'B3', 'C1', 'C2', 'C3'])

    def _record(self):
        return _rec(_lib.MED_SELLMEIER, self.coefs), (None, None)

    def __eq__(self, rhs):
        return type(rhs) == type(self) and self.coefs == rhs.coefs

    def __hash__(self):
        return hash(("batoid.SellmeierMedium", self.coefs))

    def __repr__(self):
        return f"SellmeierMedium({self.coefs})"


class SumitaMedium(Medium):
    """Sumita (Schott) dispersion, coefficients (A0, ..., A5)
    (medium.cpp:186-191)."""

    def __init__(self, *args, **kwargs):
        self.coefs = _six(args, kwargs, ['A0', 'A1', 'A2', 'A3', 'A4', 'A5'])

    def _record(self):
        return _rec(_lib.MED_SUMITA, self.coefs), (None, None)

    def __eq__(self, rhs):
        return type(rhs) == type(self) and self.coefs == rhs.coefs

    def __hash__(self):
        return hash(("batoid.SumitaMedium", self.coefs))

    def __repr__(self):
        return f"SumitaMedium({self.coefs})"


class Air(Medium):
    """Air after Filippenko (1982) / Edlen (1953); pressure and water-vapour
    pressure in kPa, temperature in K (medium.cpp:213-245)."""

    def __init__(self, pressure=69.328, temperature=293.15, h2o_pressure=1.067):
        self.pressure = pressure
        self.temperature = temperature
        self.h2o_pressure = h2o_pressure

    def _record(self):
        # the reference constructor's unit conversions (medium.cpp:229-231)
        P = self.pressure * 7.50061683
        T = self.temperature - 273.15
        W = self.h2o_pressure * 7.50061683
        return (_rec(_lib.MED_AIR, [P, T, W, self.pressure, self.temperature,
                                    self.h2o_pressure]), (None, None))

    def __eq__(self, rhs):
        return (type(rhs) == type(self) and self.pressure == rhs.pressure
                and self.temperature == rhs.temperature
                and self.h2o_pressure == rhs.h2o_pressure)

    def __hash__(self):
        return hash(("batoid.Air", self.pressure, self.temperature, self.h2o_pressure))

    def __repr__(self):
        return f"Air({self.pressure}, {self.temperature}, {self.h2o_pressure})"
