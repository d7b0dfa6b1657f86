"""The reference-side binding that INTEGRATION.md documents, executed as
written: its ``intersect`` and ``refract`` (the replacements of
``_batoid.intersect`` / ``_batoid.refract``, pysrc/batoid.cpp:55-57) must give
the same rays, bit for bit, as the package's own calls."""
import os
import re

import numpy as np
import pytest

import batoid_amd as bt
from helpers import random_rays, transforms, assert_cols_identical, SILICA

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COLS = ("x", "y", "z", "vx", "vy", "vz", "t", "wavelength", "flux", "vignetted", "failed")


def documented_binding():
    """The python block of INTEGRATION.md, executed in a fresh namespace."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    assert len(blocks) == 1, "INTEGRATION.md should hold one python block"
    ns = {}
    exec(compile(blocks[0], "INTEGRATION.md", "exec"), ns)
    return ns


def _rays(seed):
    d = random_rays(4096, seed=seed)
    return bt.RayVector(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["t"],
                        d["wavelength"], d["flux"], d["vignetted"], d["failed"])


def _host(rv):
    return {k: getattr(rv, k) for k in COLS}


@pytest.mark.parametrize("coating", [None, bt.SimpleCoating(0.02, 0.97)])
@pytest.mark.parametrize("frame", ["identity", "tilted"])
@pytest.mark.parametrize("surface", [bt.Asphere(19.835, -1.215, [0.0, -1.381e-9]),
                                     bt.Sphere(-3.1)], ids=["asphere", "sphere"])
def test_documented_intersect_is_bt_intersect(surface, frame, coating):
    ns = documented_binding()
    dr, drot = dict(transforms())[frame]
    cs = bt.CoordSys()
    target = bt.CoordSys(origin=dr, rot=drot)
    a = _rays(3)
    b = _rays(3)
    surface.intersect(a, coordSys=target, coating=coating)
    cols = {k: b.torch(k) for k in COLS}
    ns["intersect"](surface._dev.ptr(), dr, drot, cols, coating=coating)
    assert b.coordSys == cs                    # the raw call leaves frames to the caller
    assert_cols_identical(_host(b), _host(a), what=f"intersect {frame}")


@pytest.mark.parametrize("coating", [None, bt.SimpleCoating(0.04, 0.95)])
def test_documented_refract_is_bt_refract(coating):
    ns = documented_binding()
    surface = bt.Asphere(2.529, -1.57, [0.0, -0.001656])
    m1, m2 = bt.ConstMedium(1.0), bt.SellmeierMedium(*SILICA)
    dr, drot = dict(transforms())["tilted"]
    a = _rays(5)
    b = _rays(5)
    surface.refract(a, m1, m2, coordSys=bt.CoordSys(origin=dr, rot=drot), coating=coating)
    cols = {k: b.torch(k) for k in COLS}
    ns["refract"](surface._dev.ptr(), dr, drot, m2._dev.ptr(), cols, coating=coating)
    assert_cols_identical(_host(b), _host(a), what="refract")
    assert np.isfinite(b.vx).any()
