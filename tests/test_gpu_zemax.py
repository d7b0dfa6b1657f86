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


class GpuTracer:
    """The GPU library through batoid's API (Optic.traceFull)."""

    def rays(self, x, y, z, vx, vy, vz, t, w):
        return bt.RayVector(x, y, z, vx, vy, vz, t=t, wavelength=w)

    def from_stop(self, x, y, optic, stop, w, dirCos):
        return bt.RayVector.fromStop(x, y, optic=optic, stopSurface=stop, wavelength=w,
                                     dirCos=dirCos)

    def propagate(self, rv, t):
        rv.propagate(t)

    def trace_full(self, optic, rv):
        out = []
        for name, d in optic.traceFull(rv).items():
            o = d['out'].toCoordSys(bt.globalCoordSys)
            out.append((name, o.r[0], o.v[0]))
        return out


@pytest.mark.parametrize("case", ["hsc", "decam", "cbp", "comcam"])
def test_zemax_traces(case):
    """HSC, DECam, CBP and ComCam Zemax traces (tests/test_zemax.py:15-667)."""
    import zemax_cases
    getattr(zemax_cases, case)(GpuTracer())


@pytest.mark.parametrize("name", ["HSC_old.yaml", "DECam.yaml", "ComCam_r.yaml"])
def test_fused_plan_bit_exact_on_golden_optics(name):
    """The fused plan kernel on the golden optics (Baffles, tilted frames,
    table/Sumita/Sellmeier glasses, union/polygon/rectangle obscurations)
    is bit-identical to the oracle, flags included."""
    import zemax_cases
    from oracle.host import HostRays, HostPlan
    from batoid_amd.plan import PlanSpec, collect_steps
    optic = zemax_cases.optic(name)
    for thx, thy, w in [(0.0, 0.0, 500e-9), (0.004, -0.006, 800e-9)]:
        rv = bt.RayVector.asPolar(optic=optic, wavelength=w, theta_x=thx, theta_y=thy,
                                  nrad=60, naz=360)
        ref = HostRays.from_rv(rv)
        HostPlan(PlanSpec(collect_steps(optic), rv.coordSys)).trace(ref)
        optic.trace(rv)
        for k in ("x", "y", "z", "vx", "vy", "vz", "t"):
            a, b = getattr(rv, k), ref[k]
            same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
            assert same.all(), f"{name} {k}: {int((~same).sum())} rays differ"
        for k in ("vignetted", "failed"):
            np.testing.assert_array_equal(getattr(rv, k), ref[k].astype(bool))
        assert 0 < int((~rv.vignetted).sum()) < len(rv)


def test_plan_follows_reassigned_items():
    """Reassigning an item's obscuration or surface between traces (as the
    reference's own tests do) is honoured by the next fused trace."""
    from oracle.host import HostRays, HostPlan
    from batoid_amd.plan import PlanSpec, collect_steps
    telescope = bt.Optic.fromYaml(os.path.join(GOLD, "optics", "LSST_r.yaml"))

    def check():
        rv = bt.RayVector.asPolar(optic=telescope, wavelength=625e-9, theta_x=0.0,
                                  theta_y=0.0, nrad=40, naz=200, inner=0.0)
        ref = HostRays.from_rv(rv)
        HostPlan(PlanSpec(collect_steps(telescope), rv.coordSys)).trace(ref)
        telescope.trace(rv)
        for k in ("x", "y", "z", "t"):
            np.testing.assert_array_equal(getattr(rv, k), ref[k])
        np.testing.assert_array_equal(rv.vignetted, ref["vignetted"].astype(bool))
        return int(rv.vignetted.sum())

    n0 = check()
    telescope.clearObscuration()
    n1 = check()
    telescope['LSST.M1'].obscuration = bt.ObscNegation(bt.ObscAnnulus(0.61 * 4.18, 4.18))
    n2 = check()
    telescope['LSST.M2'].surface = bt.Quadric(6.788, -0.2)
    check()
    assert n1 < n2 <= n0, (n0, n1, n2)
