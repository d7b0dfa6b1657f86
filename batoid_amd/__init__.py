"""batoid_amd: MI355X-native trace + Huygens-PSF hot path of the batoid ray
tracer, with batoid's Python API (Optic.trace, RayVector, huygensPSF,
wavefront).  All ray arithmetic runs in hand-written gfx950 HIP kernels in
libbatoid_amd.so, reached through a C ABI (include/batoid_amd.h)."""
__version__ = "0.1.0"

from .coordSys import CoordSys, RotX, RotY, RotZ
from .coordTransform import CoordTransform
from .medium import Medium, ConstMedium, TableMedium, SellmeierMedium, SumitaMedium, Air
from .constants import globalCoordSys, vacuum
from .surface import (Surface, Plane, Paraboloid, Sphere, Quadric, Asphere, Bicubic,
                      Sum, Tilted, Zernike)
from .obscuration import (Obscuration, ObscCircle, ObscAnnulus, ObscRectangle, ObscRay,
                          ObscPolygon, ObscNegation, ObscUnion, ObscIntersection)
from .rayVector import RayVector, concatenateRayVectors
from .trace import (intersect, applyForwardTransform, applyReverseTransform, obscure,
                    reflect, refract, rSplit)
from .optic import (Optic, CompoundOptic, Lens, Interface, RefractiveInterface, Mirror,
                    Detector, Baffle, OPDScreen)
from .lattice import Lattice
from .coating import Coating, SimpleCoating
from . import parse
from . import utils
from . import global_vars
from .analysis import huygensPSF, wavefront, dkdu, zernike, zernikeGQ
from ._lib import BatoidAMDError
