"""Transformation between two coordinate systems (batoid/coordTransform.py).

``dr = R_from^T (o_to - o_from)`` and ``drot = R_from^T R_to``
(coordTransform.py:16-20); local = drot^T (r - dr), reverse r = drot r + dr.
The ray-vector forms run on the GPU; the array forms return host arrays."""
import numpy as np

from .coordSys import CoordSys


class CoordTransform:
    def __init__(self, fromSys, toSys):
        self.fromSys = fromSys
        self.toSys = toSys
        self.dr = fromSys.rot.T @ (toSys.origin - fromSys.origin)
        self.drot = fromSys.rot.T @ toSys.rot

    def __getstate__(self):
        return self.fromSys, self.toSys

    def __setstate__(self, d):
        self.__init__(*d)

    def __eq__(self, rhs):
        if not isinstance(rhs, CoordTransform):
            return False
        return self.fromSys == rhs.fromSys and self.toSys == rhs.toSys

    def __ne__(self, rhs):
        return not (self == rhs)

    def __hash__(self):
        return hash(("CoordTransform", self.fromSys, self.toSys))

    def __repr__(self):
        return f"CoordTransform({self.fromSys!r}, {self.toSys!r})"

    def applyForward(self, rv):
        """Transform a RayVector in place into ``toSys`` (on the GPU)."""
        from .trace import applyForwardTransform
        return applyForwardTransform(self, rv)

    def applyReverse(self, rv):
        """Transform a RayVector in place into ``fromSys`` (on the GPU)."""
        from .trace import applyReverseTransform
        return applyReverseTransform(self, rv)

    def applyForwardArray(self, x, y, z):
        """Forward transform of coordinate arrays; returns a new (3, n) array
        (coordTransform.py:70-92)."""
        r = np.array([x, y, z], dtype=float, copy=True).T
        r -= self.dr
        return self.drot.T @ r.T

    def applyReverseArray(self, x, y, z):
        """Reverse transform of coordinate arrays (coordTransform.py:94-116)."""
        r = np.array([x, y, z], dtype=float, copy=True)
        r = (self.drot @ r).T
        r += self.dr
        return r.T
