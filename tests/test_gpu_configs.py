"""The benchmark configurations themselves under test, each against the
oracle (the C restatement of the reference, itself pinned bit-for-bit to the
reference's compiled core):

* C1 -- the Newtonian telescope's 1e6 seeded rays: the fused GPU trace is
  bit-identical to the oracle in every column and flag, and the on-axis
  paraboloid images every unvignetted ray to its focus (the reference's
  tests/test_Paraboloid.py:185-202 property);
* C4 -- the TMA on a 512 x 512 pupil grid -> 256 x 256 Huygens PSF at
  0.5 um: the MFMA GEMM (FP32 and split-BF16) within 1e-5 relative L2 of the
  FP64 direct sum over the whole image, and of the oracle's direct sum on
  pixel blocks (core, shoulder and wing);
* C5 -- PSFBatch on the perturbed TMA (Sum(Asphere, Zernike, Bicubic)
  mirrors) with 128 x 128 PSFs: the batched trace bit-identical to the
  oracle, every segment's PSF within 1e-5 of the oracle's direct sum;
* the default (fftPSF) lattice of huygensPSF -- dkdu, the reciprocal lattice
  with pad 2 (analysis.py:10-56, 234-259): dkdu against a fit of the
  oracle-traced rays, the image against the obscured-Airy closed form and
  against the oracle on its own lattice;
* trace(reverse=True) -- the reference's round trip (tests/test_Optic.py:60-94)
  and bit-identity with the oracle's reverse plan.
"""
import numpy as np
import pytest
from scipy.special import j1

import batoid_amd as bt
import oracle
from oracle.host import HostRays, HostPlan, nthreads
from batoid_amd import workloads
from batoid_amd.analysis import huygens_amplitudes, _chief_index, dkdu
from batoid_amd.batch import PSFBatch
from batoid_amd.plan import PlanSpec, collect_steps
from batoid_amd.utils import fieldToDirCos
from helpers import assert_cols_identical, FCOLS, BCOLS

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def _cols(rv):
    return {k: getattr(rv, k) for k in FCOLS + BCOLS}


def _oracle_trace(optic, rv, reverse=False):
    h = HostRays.from_rv(rv)
    HostPlan(PlanSpec(collect_steps(optic, reverse), rv.coordSys)).trace(h)
    return h


def _oracle_block(rays, point, prim, n, i0, j0, m, time):
    """The oracle's FP64 direct sum on pixels [j0, j0+m) x [i0, i0+m) of the
    n x n image whose pixel (j, i) sits at point + (i - n//2) a1 + (j - n//2) a2."""
    p0 = (np.asarray(point[:2], float) + (i0 + m // 2 - n // 2) * prim[0]
          + (j0 + m // 2 - n // 2) * prim[1])
    h = HostRays.from_rv(rays)
    out = np.empty(2 * m * m)
    oracle.load().or_huygens_direct(
        h.struct, bt._lib.dvec(p0), bt._lib.dvec(prim[0]), bt._lib.dvec(prim[1]),
        m, m, float(time), 1, out.ctypes.data, nthreads())
    return out.view(np.complex128).reshape(m, m)


# ---------------------------------------------------------------------------
# C1: Newtonian, 1e6 seeded rays
# ---------------------------------------------------------------------------
def test_c1_newtonian_trace_bit_exact():
    print("case C1 newtonian 1e6 rays, seed 57721", flush=True)
    optic = workloads.newtonian()
    rv = workloads.newtonian_rays(optic)
    assert len(rv) == 1_000_000
    ref = _oracle_trace(optic, rv)
    optic.trace(rv)
    assert_cols_identical(_cols(rv), ref.cols, what="C1")
    vig = rv.vignetted
    # 0.2-1.0 m clear annulus of a uniform unit disk: 4 % of rays vignetted
    assert 0.035 < vig.mean() < 0.045
    assert not rv.failed.any()


def test_c1_on_axis_rays_meet_at_the_focus():
    """Every unvignetted on-axis ray of the C1 optic lands on the focus
    (z = R/2 is the detector), 1e6 rays."""
    print("case C1 on-axis focus", flush=True)
    optic = workloads.newtonian()
    rv = bt.RayVector.asPolar(optic=optic, outer=1.0, inner=0.0, nrandom=1_000_000,
                              rng=np.random.default_rng(57721), wavelength=500e-9,
                              theta_x=0.0, theta_y=0.0)
    optic.trace(rv)
    w = ~rv.vignetted
    np.testing.assert_allclose(rv.x[w], 0.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(rv.y[w], 0.0, rtol=0, atol=1e-12)


def test_paraboloid_reflects_to_focus():
    """tests/test_Paraboloid.py:185-202 of the reference, on the GPU."""
    rng = np.random.default_rng(5772156)
    size = 10_000
    for _ in range(20):
        R = rng.normal(0, 3.0)
        para = bt.Paraboloid(R)
        x = rng.normal(size=size)
        y = rng.normal(size=size)
        rv = bt.RayVector(x, y, -1000, 0, 0, 1)
        para.reflect(rv)
        t = rv.t[0] - rv.r[0, 0] / rv.v[0, 0]
        focus = rv.positionAtTime(t)
        np.testing.assert_allclose(focus[:, 0], 0, atol=1e-12)
        np.testing.assert_allclose(focus[:, 1], 0, atol=1e-12)
        np.testing.assert_allclose(focus[:, 2], R / 2, atol=1e-12)


# ---------------------------------------------------------------------------
# C4: TMA 512^2 pupil -> 256^2 PSF
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4_rays():
    optic = workloads.tma()
    thx, thy = workloads.tma_fields()
    rays = bt.RayVector.asGrid(optic=optic, wavelength=620e-9,
                               dirCos=fieldToDirCos(thx[0], thy[0]), nx=512)
    optic.trace(rays)
    return rays


def test_c4_trace_bit_exact(c4_rays):
    print("case C4 trace 512^2 grid", flush=True)
    optic = workloads.tma()
    thx, thy = workloads.tma_fields()
    rv = bt.RayVector.asGrid(optic=optic, wavelength=620e-9,
                             dirCos=fieldToDirCos(thx[0], thy[0]), nx=512)
    ref = _oracle_trace(optic, rv)
    assert_cols_identical(_cols(c4_rays), ref.cols, what="C4 trace")
    nsum = int((~c4_rays.vignetted & ~c4_rays.failed).sum())
    assert 1.2e5 < nsum < 1.3e5


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_c4_psf_gemm_vs_fp64(c4_rays, precision):
    print(f"case C4 PSF 256^2 {precision}", flush=True)
    rays = c4_rays
    point = rays.r[_chief_index(512)]
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    time = float(rays.t[0])
    n = 256
    direct = huygens_amplitudes(rays, point, prim, n, n, time, method='direct')
    got = huygens_amplitudes(rays, point, prim, n, n, time, method='gemm',
                             precision=precision)
    err = rel_l2(got, direct)
    print(f"C4 {precision}: rel L2 vs FP64 direct (whole image) {err:.2e}")
    assert err < 1e-5
    # the oracle on three 32 x 32 blocks: core, shoulder, wing
    peak = np.unravel_index(np.argmax(np.abs(direct)), direct.shape)
    blocks = [(max(0, min(n - 32, peak[1] - 16)), max(0, min(n - 32, peak[0] - 16))),
              (150, 40), (0, 200)]
    ref = np.concatenate([_oracle_block(rays, point, prim, n, i0, j0, 32, time).ravel()
                          for i0, j0 in blocks])
    dev = np.concatenate([direct[j0:j0 + 32, i0:i0 + 32].ravel() for i0, j0 in blocks])
    mm = np.concatenate([got[j0:j0 + 32, i0:i0 + 32].ravel() for i0, j0 in blocks])
    # the blocks' lattice origins are recomputed on the host: a pixel 1 ulp
    # away at |p| ~ 0.1 m shifts a phase by ~1e-10 rad, hence 1e-9 (the two
    # FP64 sums on the same lattice agree to 1e-12 in test_gpu_huygens)
    assert rel_l2(dev, ref) < 1e-9
    err_o = rel_l2(mm, ref)
    print(f"C4 {precision}: rel L2 vs oracle (3 blocks) {err_o:.2e}")
    assert err_o < 1e-5


# ---------------------------------------------------------------------------
# C5: batch on the perturbed TMA, 128^2 PSFs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_c5_perturbed_batch_vs_oracle(precision):
    print(f"case C5 perturbed batch 128^2 {precision}", flush=True)
    optic = workloads.perturbed_tma()
    thx, thy = workloads.tma_fields(n=2, seed=1)
    fx = np.repeat(thx, 2)
    fy = np.repeat(thy, 2)
    wl = np.array([490e-9, 1050e-9, 630e-9, 910e-9])
    nxs = [48, 160, 97, 64]
    b = PSFBatch(optic, fx, fy, wl, nxs, nxOut=128, dx=0.5e-6, precision=precision,
                 chunk=1024, nwg=96)
    b.run()
    # the batched trace is the oracle's trace of the same rays
    ref = HostRays.from_rv(b.rays_in)
    HostPlan(PlanSpec(collect_steps(optic), b.rays_in.coordSys)).trace(ref)
    assert_cols_identical(_cols(b.rays), ref.cols, what="C5 trace")
    amp = b.amplitudes()
    for s in range(b.nseg):
        lo, hi = int(b.bounds[s]), int(b.bounds[s + 1])
        rays = b.rays[lo:hi]
        point = rays.r[_chief_index(nxs[s])]
        ora = _oracle_block(rays, point, b.lattices[s], 128, 0, 0, 128, float(rays.t[0]))
        err = rel_l2(amp[s], ora)
        nsum = int((~rays.vignetted & ~rays.failed).sum())
        print(f"C5 segment {s} nx={nxs[s]} ({nsum} rays): rel L2 vs oracle {err:.2e}")
        assert nsum > 0
        assert err < 1e-5


# ---------------------------------------------------------------------------
# the default (fftPSF) lattice of huygensPSF
# ---------------------------------------------------------------------------
def test_dkdu_matches_oracle_traced_fit():
    """dkdu (analysis.py:17-56): the GPU's fit equals a numpy fit of the same
    pupil rays traced by the oracle."""
    optic = workloads.tma()
    thx, thy = np.deg2rad(0.3), np.deg2rad(-0.2)
    got = dkdu(optic, thx, thy, 620e-9)
    rays = bt.RayVector.asPolar(optic=optic, theta_x=thx, theta_y=thy, wavelength=620e-9,
                                nrad=6, naz=36)
    ux, uy = rays.x.copy(), rays.y.copy()
    h = _oracle_trace(optic, rays)
    w = ~h["vignetted"].astype(bool)
    v = np.stack([h["vx"], h["vy"], h["vz"]], axis=1)
    # k = 2 pi n vhat / lambda with n = 1/|v|: 2 pi v / (lambda |v|^2)
    k = 2 * np.pi * v / (h["wavelength"][:, None] * np.sum(v * v, axis=1)[:, None])
    A = np.stack([np.ones(w.sum()), ux[w], uy[w]], axis=1)
    sol_x = np.linalg.lstsq(A, k[w, 0], rcond=None)[0]
    sol_y = np.linalg.lstsq(A, k[w, 1], rcond=None)[0]
    expect = np.array([[sol_x[1], sol_y[1]], [sol_x[2], sol_y[2]]])
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=0)
    # the focal-plane scale: dk/du ~ -2 pi / (lambda f_eff), f_eff ~ 10.3 m
    assert -2 * np.pi / (620e-9 * 9.0) < got[0, 0] < -2 * np.pi / (620e-9 * 12.0)


def _slow_paraboloid(focal=1.5, diam=0.3, obscuration=0.25):
    """The reference's tests/test_analysis.py:217-246 telescope: one
    Paraboloid mirror (R = 2 f) with an annular clear aperture, a flat
    detector at the focus, in vacuum."""
    from batoid_amd.parse import parse_optic
    return parse_optic({
        'type': 'CompoundOptic', 'name': 'slow', 'inMedium': 1.0, 'backDist': 10.0,
        'pupilSize': diam, 'pupilObscuration': obscuration,
        'items': [
            {'type': 'Mirror', 'name': 'Mirror', 'surface': {'type': 'Paraboloid',
                                                             'R': 2 * focal},
             'obscuration': {'type': 'ClearAnnulus', 'inner': 0.5 * obscuration * diam,
                             'outer': 0.5 * diam}},
            {'type': 'Detector', 'name': 'detector', 'surface': {'type': 'Plane'},
             'coordSys': {'z': focal}}]})


@pytest.mark.parametrize("method", ["gemm", "direct"])
def test_default_lattice_psf_is_the_obscured_airy(method):
    """On axis, an f/5 paraboloid's default-lattice (fftPSF) PSF is the
    obscured Airy pattern [2 J1(x)/x - e^2 2 J1(e x)/(e x)]^2/(1-e^2)^2,
    x = pi D r/(lambda f), e = 0.25 (replaces galsim.Airy of the reference's
    tests/test_analysis.py:216-357), on the lattice 2 pi/(pad nx du dk/du)
    with dk/du = -2 pi/(lambda f) and pad 2."""
    focal, diam, eps, lam, nx = 1.5, 0.3, 0.25, 500e-9, 96
    optic = _slow_paraboloid(focal, diam, eps)
    psf = bt.huygensPSF(optic, 0.0, 0.0, lam, nx=nx, method=method)
    assert psf.array.shape == (2 * nx, 2 * nx)          # pad factor 2
    du = diam / (nx - 2)
    scale = lam * focal / (2 * nx * du)
    prim = psf.primitiveVectors
    np.testing.assert_allclose(np.abs(np.diag(prim)), [scale, scale], rtol=2e-3)
    # off-diagonal terms come only from the fit's sampling (~1e-5 relative)
    assert np.abs(prim[0, 1]) < 1e-4 * scale and np.abs(prim[1, 0]) < 1e-4 * scale
    I = psf.array / psf.array.max()
    xy = psf.coords                                     # [i, j] -> (x, y); image is [j, i]
    r = np.hypot(xy[..., 0].T, xy[..., 1].T)
    u = np.pi * diam * r / (lam * focal)
    with np.errstate(invalid='ignore', divide='ignore'):
        model = np.where(u == 0, 1 - eps ** 2,
                         2 * j1(u) / u - eps ** 2 * 2 * j1(eps * u) / (eps * u))
    model = model ** 2 / (1 - eps ** 2) ** 2
    core = u < 12.0                                     # the first three rings
    np.testing.assert_allclose(I[core], model[core], rtol=0, atol=1e-2)
    # the chief ray meets the focus: the peak is the lattice centre
    assert np.unravel_index(np.argmax(I), I.shape) == (nx, nx)


def test_default_lattice_psf_vs_oracle_and_explicit_form():
    """Off axis through the TMA: the default-lattice image equals the
    explicit-dx call on the same primitive vectors, and the oracle's direct
    sum on that lattice (the sheared lattice exercises the reference's
    unsigned b1 = (a2[1], a2[0]) and the [y, x] transposition)."""
    optic = workloads.tma()
    thx, thy = np.deg2rad(0.9), np.deg2rad(0.6)
    lam, nx = 620e-9, 48
    psf = bt.huygensPSF(optic, thx, thy, lam, nx=nx, method='direct')
    prim = psf.primitiveVectors
    assert prim[0, 1] != 0.0 and prim[1, 0] != 0.0      # sheared
    again = bt.huygensPSF(optic, thx, thy, lam, nx=nx, dx=prim[0], dy=prim[1],
                          nxOut=2 * nx, method='direct')
    np.testing.assert_array_equal(again.array, psf.array)
    rays = bt.RayVector.asGrid(optic=optic, wavelength=lam, dirCos=fieldToDirCos(thx, thy),
                               nx=nx)
    optic.trace(rays)
    point = rays.r[_chief_index(nx)]
    ora = _oracle_block(rays, point, prim, 2 * nx, 0, 0, 2 * nx, float(rays.t[0]))
    assert rel_l2(psf.array, np.abs(ora) ** 2) < 1e-12
    gemm = bt.huygensPSF(optic, thx, thy, lam, nx=nx)
    assert rel_l2(gemm.array, psf.array) < 2e-5


# ---------------------------------------------------------------------------
# reverse trace
# ---------------------------------------------------------------------------
def _hsc():
    import os
    return bt.Optic.fromYaml(os.path.join(os.path.dirname(__file__), "golden", "optics",
                                          "HSC.yaml"))


def test_reverse_trace_round_trip():
    """tests/test_Optic.py:60-94 of the reference: trace forward, turn the
    rays around, trace with reverse=True, and land back on the inputs."""
    telescope = _hsc()
    init_rays = bt.RayVector.asGrid(backDist=25, lx=8.3, nx=64, theta_x=0.005,
                                    theta_y=0.005, wavelength=500e-9,
                                    medium=bt.ConstMedium(1.0))
    forward_rays = telescope.trace(init_rays.copy())
    forward_rays.propagate(40.0)
    reverse_rays = forward_rays.copy()
    for k in ("vx", "vy", "vz", "t"):
        col = reverse_rays.torch(k)
        col.mul_(-1)
    final_rays = telescope.trace(reverse_rays.copy(), reverse=True)
    final_rays.propagate(0.0)
    final_rays.toCoordSys(bt.CoordSys())
    w = ~final_rays.vignetted
    assert w.sum() > 1000
    np.testing.assert_allclose(init_rays.x[w], final_rays.x[w])
    np.testing.assert_allclose(init_rays.y[w], final_rays.y[w])
    np.testing.assert_allclose(init_rays.z[w], final_rays.z[w])
    np.testing.assert_allclose(init_rays.vx[w], -final_rays.vx[w])
    np.testing.assert_allclose(init_rays.vy[w], -final_rays.vy[w])
    np.testing.assert_allclose(init_rays.vz[w], -final_rays.vz[w])
    np.testing.assert_allclose(final_rays.t[w], 0, atol=1e-9)


def test_reverse_trace_bit_exact():
    """The fused reverse plan equals the oracle's reverse plan and the
    item-by-item reverse trace (optic.py:1013-1049 with reverse=True)."""
    telescope = _hsc()
    rays = bt.RayVector.asGrid(backDist=25, lx=8.3, nx=96, theta_x=0.005, theta_y=0.005,
                               wavelength=500e-9, medium=bt.ConstMedium(1.0))
    telescope.trace(rays)
    rays.propagate(40.0)
    for k in ("vx", "vy", "vz", "t"):
        rays.torch(k).mul_(-1)
    ref = _oracle_trace(telescope, rays, reverse=True)
    step = rays.copy()
    fused = telescope.trace(rays.copy(), reverse=True)
    assert_cols_identical(_cols(fused), ref.cols, what="reverse fused")
    for iface, _, _ in collect_steps(telescope, reverse=True):
        iface.trace(step, reverse=True)
    assert_cols_identical(_cols(step), ref.cols, what="reverse per item")
    assert fused.coordSys == step.coordSys
