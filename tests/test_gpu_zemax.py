"""Golden vectors through the full GPU API: Zemax ray traces through LSST
(reference tests/test_zemax.py:412-462, 1e-10 m) and Zemax wavefront maps
(tests/test_zemax.py:223-269, 1e-11 m)."""
import json
import os

import numpy as np
import pytest

import batoid_amd as bt

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def lsst():
    return bt.Optic.fromYaml(os.path.join(GOLD, "optics", "LSST_g_500.yaml"))


def test_lsst_trace_vs_zemax():
    d = json.load(open(os.path.join(GOLD, "lsst_trace.json")))
    telescope = lsst()
    for tr in d["traces"]:
        arr = np.array(tr["rows"])
        ray = bt.RayVector.fromStop(
            tr["Px"] * 4.18, tr["Py"] * 4.18, optic=telescope, wavelength=500e-9,
            theta_x=np.deg2rad(tr["Hx"] * 1.75), theta_y=np.deg2rad(tr["Hy"] * 1.75),
            projection='zemax')
        tf = telescope.traceFull(ray)
        for surface, iz in zip(tf.values(), d["zemax_surfaces"]):
            r = surface['out'].toCoordSys(bt.globalCoordSys)
            n = 1. / np.sqrt(np.sum(r.v ** 2))
            np.testing.assert_allclose(r.x, arr[iz][0] / 1e3, rtol=0, atol=1e-10)
            np.testing.assert_allclose(r.y, arr[iz][1] / 1e3, rtol=0, atol=1e-10)
            np.testing.assert_allclose(np.abs(r.z), np.abs(arr[iz][2] / 1e3), rtol=0, atol=1e-10)
            np.testing.assert_allclose(np.abs(r.vx * n), np.abs(arr[iz][3]), rtol=0, atol=1e-10)
            np.testing.assert_allclose(np.abs(r.vy * n), np.abs(arr[iz][4]), rtol=0, atol=1e-10)
            np.testing.assert_allclose(np.abs(r.vz * n), np.abs(arr[iz][5]), rtol=0, atol=1e-10)


def test_fused_trace_equals_traceFull_end():
    telescope = lsst()
    rays = bt.RayVector.asPolar(optic=telescope, wavelength=620e-9, theta_x=0.01,
                                theta_y=-0.005, nrad=30, naz=120)
    r2 = rays.copy()
    telescope.trace(rays)
    out = list(telescope.traceFull(r2).values())[-1]['out']
    assert rays == out or np.array_equal(rays.x, out.x)
    for k in ("x", "y", "z", "t", "vignetted", "failed"):
        np.testing.assert_array_equal(getattr(rays, k), getattr(out, k))


def test_lsst_wavefront_vs_zemax():
    d = np.load(os.path.join(GOLD, "lsst_wf.npz"))
    telescope = lsst()
    for thx, thy, Zwf in zip(d["thx_deg"], d["thy_deg"], d["maps"]):
        Zwf = Zwf[::-1]
        bwf = bt.wavefront(telescope, np.deg2rad(thx), np.deg2rad(thy), 500e-9, nx=32,
                           reference='chief', projection='zemax')
        Zwf = np.ma.MaskedArray(data=Zwf, mask=Zwf == 0)
        np.testing.assert_allclose(Zwf * 500e-9, bwf.array * 500e-9, atol=1e-11, rtol=0)
