"""Coatings: reflected / transmitted flux fractions at an Interface
(batoid/coating.py; src/coating.cpp).  The reference has one kind,
SimpleCoating, constant in wavelength and incidence angle; on the GPU it is
a two-number record (bt_coating) passed with the launch."""
from . import _lib


class Coating:
    """Base class (coating.py:5-77)."""

    def getCoefs(self, wavelength, cosIncidenceAngle):
        return self.getReflect(wavelength, cosIncidenceAngle), \
            self.getTransmit(wavelength, cosIncidenceAngle)

    def __ne__(self, rhs):
        return not (self == rhs)


class SimpleCoating(Coating):
    """Reflectivity and transmissivity constant with wavelength and incidence
    angle (coating.py:80-114, coating.cpp:10-39)."""

    def __init__(self, reflectivity, transmissivity):
        self.reflectivity = reflectivity
        self.transmissivity = transmissivity

    def getReflect(self, wavelength, cosIncidenceAngle):
        return self.reflectivity

    def getTransmit(self, wavelength, cosIncidenceAngle):
        return self.transmissivity

    @property
    def _record(self):
        return _lib.BtCoating(float(self.reflectivity), float(self.transmissivity))

    def __eq__(self, rhs):
        return (isinstance(rhs, SimpleCoating) and self.reflectivity == rhs.reflectivity
                and self.transmissivity == rhs.transmissivity)

    def __hash__(self):
        return hash(("SimpleCoating", self.reflectivity, self.transmissivity))

    def __getstate__(self):
        return self.reflectivity, self.transmissivity

    def __setstate__(self, args):
        self.__init__(*args)

    def __repr__(self):
        return f"SimpleCoating({self.reflectivity}, {self.transmissivity})"
