"""GPU parity of bt_generate_rays -- RayVector.asGrid / asPolar / asSpokes /
fromStop -- against the oracle's restatement of the reference generators
(oracle/raygen.py) on the same arguments.

Bar: lattice, random-lattice and given-point rays are bit-identical in all
11 columns (the stop coordinates are the reference's exact numpy
expressions and the rest is FP64 with no contraction on either side).
Hexapolar and random-annulus rays take (r cos th, r sin th) with the
device's FP64 sin/cos, which may differ from numpy's in the last bit or
two: their positions are held to 2e-14 m (a few ulps at the 10-40 m
source-plane distances) and every other column stays identical; traced
back to the stop plane they sit at their ring's radius to 2e-14 m."""
import os

import numpy as np
import pytest
import torch

import batoid_amd as bt
import oracle
from oracle import raygen as G
from oracle.host import host_records
from batoid_amd import workloads
from helpers import FCOLS, BCOLS

pytestmark = pytest.mark.gpu

POS_TOL = 2e-14


def _host(rv):
    torch.cuda.synchronize()
    return {k: getattr(rv, k) for k in FCOLS + BCOLS}


def _oracle_n(medium, w):
    tab = host_records(medium)
    w = np.ascontiguousarray([w], dtype=float)
    out = np.empty(1)
    oracle.load().or_get_n_n(tab.ptr(), w.ctypes.data, out.ctypes.data, 1, 1)
    return out[0]


def _oracle_rays(x, y, stop, coordSys, medium, wavelength, dirCos, backDist, flux=1.0):
    ct = bt.CoordTransform(stop.coordSys, coordSys)
    h = G.finish(x, y, stop.surface, (ct.dr, ct.drot), dirCos, _oracle_n(medium, wavelength),
                 backDist, wavelength, flux)
    return {k: h[k] for k in FCOLS + BCOLS}


def _identical(a, b, cols=FCOLS + BCOLS, what=""):
    for k in cols:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=f"{what} {k}")


def _close_positions(a, b, what=""):
    for k in ("x", "y", "z"):
        d = np.max(np.abs(a[k] - b[k])) if len(a[k]) else 0.0
        assert d <= POS_TOL, f"{what} {k}: max |diff| {d:.3e} m"
    _identical(a, b, cols=("vx", "vy", "vz", "t", "wavelength", "flux") + BCOLS, what=what)


def _tma():
    return workloads.tma()


def _tilted_stop():
    """A curved, shifted and tilted stop: exercises the sag and the frame change."""
    cs = bt.CoordSys(origin=[0.01, -0.02, 0.35], rot=bt.RotX(0.02) @ bt.RotY(-0.01))
    return bt.Interface(bt.Sphere(-25.0), coordSys=cs)


@pytest.mark.parametrize("nx,ny", [(64, 64), (63, 63), (1, 1), (2, 2), (40, 17), (512, 512)])
def test_grid_bit_exact(nx, ny):
    optic = _tma()
    w = 620e-9
    dirCos = bt.utils.fieldToDirCos(0.004, -0.007)
    rv = bt.RayVector.asGrid(optic=optic, wavelength=w, dirCos=dirCos, nx=nx, ny=ny,
                             lx=optic.pupilSize)
    lx = optic.pupilSize
    uu, vv = G.grid_uv(nx, ny)
    x, y = G.lattice_xy((lx, 0.0), (0.0, lx), uu, vv)
    ref = _oracle_rays(x, y, optic.stopSurface, optic.coordSys, optic.inMedium, w, dirCos,
                       optic.backDist)
    assert len(rv) == nx * ny
    _identical(_host(rv), ref, what=f"grid {nx}x{ny}")


def test_grid_curved_tilted_stop_and_medium():
    stop = _tilted_stop()
    medium = bt.SellmeierMedium(0.6961663, 0.4079426, 0.8974794, 0.00467914825849,
                                0.013512063073959999, 97.93400253792099)
    w = 700e-9
    dirCos = bt.utils.fieldToDirCos(-0.01, 0.02)
    for lx, ly in [(3.0, None), ((2.0, 0.0), (0.0, 2.5))]:
        rv = bt.RayVector.asGrid(backDist=12.0, medium=medium, stopSurface=stop, wavelength=w,
                                 dirCos=dirCos, nx=48, ny=50, lx=lx, ly=ly, flux=0.25)
        lxv = (lx, 0.0) if np.isscalar(lx) else lx
        lyv = (0.0, lx) if ly is None else ly
        uu, vv = G.grid_uv(48, 50)
        x, y = G.lattice_xy(lxv, lyv, uu, vv)
        ref = _oracle_rays(x, y, stop, bt.globalCoordSys, medium, w, dirCos, 12.0, 0.25)
        _identical(_host(rv), ref, what=f"tilted stop lx={lx}")


def test_grid_random_bit_exact_and_rng_state():
    optic = _tma()
    w = 500e-9
    dirCos = bt.utils.fieldToDirCos(0.01, 0.003)
    n = 200_000
    rng = np.random.default_rng(57721)
    rv = bt.RayVector.asGrid(optic=optic, wavelength=w, dirCos=dirCos, nx=1, lx=optic.pupilSize,
                             nrandom=n, rng=rng)
    ref_rng = np.random.default_rng(57721)
    uu = ref_rng.uniform(-0.5, 0.5, size=n)
    vv = ref_rng.uniform(-0.5, 0.5, size=n)
    lx = optic.pupilSize
    x, y = G.lattice_xy((lx, 0.0), (0.0, lx), uu, vv)
    ref = _oracle_rays(x, y, optic.stopSurface, optic.coordSys, optic.inMedium, w, dirCos,
                       optic.backDist)
    _identical(_host(rv), ref, what="random grid")
    # the caller's generator moved exactly as numpy's would
    assert rng.bit_generator.state == ref_rng.bit_generator.state
    # seeded by integer: same rays (tests/test_RayVector.py:491-508)
    rv2 = bt.RayVector.asGrid(optic=optic, wavelength=w, dirCos=dirCos, nx=1,
                              lx=optic.pupilSize, nrandom=n, rng=57721)
    _identical(_host(rv2), ref, what="random grid, int seed")


def test_random_with_other_bit_generator():
    """A non-PCG64 generator: its numpy draws are used as given points."""
    stop = bt.Interface(bt.Plane())
    w = 600e-9
    dirCos = np.array([0.0, 0.0, -1.0])
    rng = np.random.Generator(np.random.MT19937(11))
    rv = bt.RayVector.asPolar(backDist=10.0, stopSurface=stop, wavelength=w, dirCos=dirCos,
                              inner=0.1, outer=0.2, nrandom=5000, rng=rng)
    ref_rng = np.random.Generator(np.random.MT19937(11))
    rr = np.sqrt(ref_rng.uniform(0.1**2, 0.2**2, size=5000))
    th = ref_rng.uniform(0, 2 * np.pi, size=5000)
    x, y = G.polar_xy(rr, th)
    ref = _oracle_rays(x, y, stop, bt.globalCoordSys, bt.vacuum, w, dirCos, 10.0)
    _identical(_host(rv), ref, what="MT19937 polar")


@pytest.mark.parametrize("nrad,naz,inner", [(40, 400, 2.558), (17, 96, 0.0), (1, 30, 0.0),
                                            (1, 30, 1.0), (300, 3000, 2.558)])
def test_polar_matches_oracle(nrad, naz, inner):
    optic = _tma()
    w = 770e-9
    dirCos = bt.utils.fieldToDirCos(-0.012, 0.005)
    rv = bt.RayVector.asPolar(optic=optic, wavelength=w, dirCos=dirCos, outer=4.18,
                              inner=inner, nrad=nrad, naz=naz)
    rr, th = G.hexapolar(4.18, inner, nrad, naz)
    x, y = G.polar_xy(rr, th)
    ref = _oracle_rays(x, y, optic.stopSurface, optic.coordSys, optic.inMedium, w, dirCos,
                       optic.backDist)
    got = _host(rv)
    assert len(got["x"]) == len(rr)
    _close_positions(got, ref, what=f"polar {nrad}x{naz} inner={inner}")
    # on the stop plane the radii and azimuths are the table's
    stop_cs = optic.stopSurface.coordSys
    srv = rv.toCoordSys(stop_cs)
    bt.Plane().intersect(srv)
    h = _host(srv)
    np.testing.assert_allclose(np.hypot(h["x"], h["y"]), rr, rtol=0, atol=2e-14)
    if inner == 0.0:
        assert abs(h["x"][-1]) < 1e-14 and abs(h["y"][-1]) < 1e-14


def test_polar_random_matches_numpy_stream():
    """C1's pupil: asPolar(nrandom=1e6, rng=default_rng(57721))."""
    optic = workloads.newtonian()
    w = 500e-9
    n = 1_000_000
    rng = np.random.default_rng(57721)
    rv = bt.RayVector.asPolar(optic=optic, wavelength=w, theta_x=np.deg2rad(0.1), theta_y=0.0,
                              outer=1.0, inner=0.0, nrandom=n, rng=rng)
    ref_rng = np.random.default_rng(57721)
    rr = np.sqrt(ref_rng.uniform(0.0**2, 1.0**2, size=n))
    th = ref_rng.uniform(0, 2 * np.pi, size=n)
    x, y = G.polar_xy(rr, th)
    dirCos = bt.utils.fieldToDirCos(np.deg2rad(0.1), 0.0)
    # no stop surface on the optic: a Plane at the global origin (rayVector.py:596-597)
    ref = _oracle_rays(x, y, bt.Interface(bt.Plane()), optic.coordSys, optic.inMedium, w,
                       dirCos, optic.backDist)
    _close_positions(_host(rv), ref, what="random polar")
    assert rng.bit_generator.state == ref_rng.bit_generator.state


def test_from_stop_and_spokes_bit_exact():
    optic = bt.Optic.fromYaml(os.path.join(os.path.dirname(__file__), "golden", "optics",
                                           "LSST_g_500.yaml"))
    rng = np.random.default_rng(31)
    x = rng.uniform(-4.0, 4.0, size=3001)
    y = rng.uniform(-4.0, 4.0, size=3001)
    w = 500e-9
    for proj in ("postel", "zemax"):
        rv = bt.RayVector.fromStop(x, y, optic=optic, wavelength=w, theta_x=0.01,
                                   theta_y=-0.005, projection=proj)
        dirCos = bt.utils.fieldToDirCos(0.01, -0.005, projection=proj)
        ref = _oracle_rays(x, y, optic.stopSurface, optic.coordSys, optic.inMedium, w, dirCos,
                           optic.backDist)
        _identical(_host(rv), ref, what=f"fromStop {proj}")
    # scalar point
    rv = bt.RayVector.fromStop(0.3, -0.2, optic=optic, wavelength=w, theta_x=0.0, theta_y=0.0)
    assert len(rv) == 1
    # spokes with Gauss-quadrature fluxes (per-ray flux column)
    rv = bt.RayVector.asSpokes(optic=optic, wavelength=w, theta_x=0.002, theta_y=0.0,
                               inner=2.558, outer=4.18, rings=6, spacing='GQ')
    Li, wq = np.polynomial.legendre.leggauss(6)
    eps = 2.558 / 4.18
    rings = np.sqrt(eps**2 + (1 + Li) * (1 - eps**2) / 2) * 4.18
    area = np.pi * (1 - eps**2)
    flux = wq * area / (2 * 13)                  # rayVector.py:790-794
    spokes = np.linspace(0, 2 * np.pi, 13, endpoint=False)
    rings, spokes = np.meshgrid(rings, spokes)
    flux = np.broadcast_to(flux, rings.shape).ravel()
    xs, ys = rings.ravel() * np.cos(spokes.ravel()), rings.ravel() * np.sin(spokes.ravel())
    ref = _oracle_rays(xs, ys, optic.stopSurface, optic.coordSys, optic.inMedium, w,
                       bt.utils.fieldToDirCos(0.002, 0.0), optic.backDist, flux)
    _identical(_host(rv), ref, what="spokes GQ")


def test_reference_asgrid_properties():
    """tests/test_RayVector.py:368-431 on the GPU's rays."""
    rng = np.random.default_rng(5772156)
    for _ in range(6):
        backDist = rng.uniform(9.0, 11.0)
        wavelength = rng.uniform(300e-9, 1100e-9)
        nx = int(rng.integers(10, 21))
        lx = rng.uniform(1.0, 10.0)
        dx = lx / (nx - 2) if nx % 2 == 0 else lx / (nx - 1)
        dirCos = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1),
                           rng.uniform(-1.2, -0.8)])
        dirCos /= np.sqrt(np.dot(dirCos, dirCos))
        g1 = bt.RayVector.asGrid(backDist=backDist, wavelength=wavelength, nx=nx, lx=lx,
                                 dirCos=dirCos)
        g2 = bt.RayVector.asGrid(backDist=backDist, wavelength=wavelength, nx=nx, dx=dx,
                                 dirCos=dirCos)
        g4 = bt.RayVector.asGrid(backDist=backDist, wavelength=wavelength, nx=nx,
                                 lx=(lx, 0.0), dirCos=dirCos)
        for g in (g2, g4):
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                np.testing.assert_allclose(getattr(g, k), getattr(g1, k), rtol=0, atol=1e-14)
        cr = (nx // 2) * nx + nx // 2 if nx % 2 == 0 else (nx * nx - 1) // 2
        np.testing.assert_allclose(np.sqrt(np.dot(g1.r[cr], g1.r[cr])), backDist)
        np.testing.assert_allclose(g1.vx, dirCos[0])
        assert not g1.vignetted.any() and not g1.failed.any()
        bt.Plane().intersect(g1)
        np.testing.assert_allclose(np.diff(g1.x)[0], dx)
        np.testing.assert_allclose(np.diff(g1.y)[0], 0, atol=1e-14)
        np.testing.assert_allclose(np.diff(g1.x)[nx - 1], -dx * (nx - 1))
        np.testing.assert_allclose(np.diff(g1.y)[nx - 1], dx)


def test_c2_pupil_full_size_properties():
    """The 9,998,544-ray C2 pupil: count, every ray on the source plane
    through the origin perpendicular to v, and back on the stop plane at
    its ring's radius (size-independent checks)."""
    optic = _tma()
    rv = workloads.tma_rays(optic, 0.005, -0.01, 630e-9)
    assert len(rv) == 9_998_544
    v = np.array([rv.torch("vx")[0].item(), rv.torch("vy")[0].item(), rv.torch("vz")[0].item()])
    vhat = v / np.linalg.norm(v)
    r = torch.stack([rv.torch("x"), rv.torch("y"), rv.torch("z")], dim=1)
    vt = torch.tensor(vhat, dtype=torch.float64, device=r.device)
    # the source plane is at distance backDist before the origin along v
    plane = (r @ vt).cpu().numpy()
    np.testing.assert_allclose(plane, -optic.backDist, rtol=0, atol=1e-12)
    srv = rv.toCoordSys(optic.stopSurface.coordSys)
    bt.Plane().intersect(srv)
    rad = torch.hypot(srv.torch("x"), srv.torch("y")).cpu().numpy()
    rr, _ = G.hexapolar(4.18, 2.558, 1120, 11080)
    np.testing.assert_allclose(rad, rr, rtol=0, atol=1e-13)
    assert not bool(srv.torch("failed").any())
