"""GPU Huygens PSF: the FP64 direct sum against the oracle's, and the MFMA
GEMM (FP32 and split-BF16) against the FP64 direct sum.

Tolerances (relative L2 of the complex amplitude image, which bounds the
intensity image's):
  direct FP64 vs oracle FP64      <= 1e-12  (device vs libm sin/cos only)
  GEMM FP32 vs FP64 direct        <= 1e-5
  GEMM split-BF16 vs FP64 direct  <= 1e-5
"""
import numpy as np
import pytest
import torch

import batoid_amd as bt
import oracle
from oracle.host import HostRays, nthreads
from batoid_amd import workloads
from batoid_amd.analysis import huygens_amplitudes

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def _traced(nx=64, theta=0.1):
    optic = workloads.newtonian()
    rays = bt.RayVector.asGrid(optic=optic, wavelength=500e-9, nx=nx,
                               dirCos=bt.utils.fieldToDirCos(np.deg2rad(theta), 0.0))
    optic.trace(rays)
    return optic, rays


def _oracle_direct(rays, point, prim, nx, ny, time):
    h = HostRays.from_rv(rays)
    out = np.empty(2 * nx * ny)
    oracle.load().or_huygens_direct(
        h.struct, bt._lib.dvec(point[:2]), bt._lib.dvec(prim[0]), bt._lib.dvec(prim[1]),
        nx, ny, float(time), 1, out.ctypes.data, nthreads())
    return out.view(np.complex128).reshape(ny, nx)


@pytest.mark.parametrize("nx,ny", [(32, 32), (48, 40), (130, 20)])
def test_direct_matches_oracle(nx, ny):
    optic, rays = _traced()
    point = rays.r[np.flatnonzero(~rays.vignetted)[0]]
    prim = np.array([[0.7e-6, 0.05e-6], [-0.03e-6, 0.6e-6]])
    time = float(rays.t[0])
    ref = _oracle_direct(rays, point, prim, nx, ny, time)
    got = huygens_amplitudes(rays, point, prim, nx, ny, time, method='direct')
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("nx,ny", [(64, 64), (48, 40), (130, 20), (256, 256)])
def test_gemm_matches_fp64(precision, nx, ny):
    print(f"case gemm {precision} {nx}x{ny}", flush=True)
    optic, rays = _traced(nx=96)
    point = np.mean(rays.r[~rays.vignetted], axis=0)
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    time = float(rays.t[0])
    ref = huygens_amplitudes(rays, point, prim, nx, ny, time, method='direct')
    got = huygens_amplitudes(rays, point, prim, nx, ny, time, method='gemm',
                             precision=precision)
    err = rel_l2(got, ref)
    print(f"rel L2 {precision} {nx}x{ny}: {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("nkeep", [1, 5, 13, 700])
def test_gemm_few_rays(precision, nkeep):
    """K splits of fewer rays than the kernels' pipeline depth or ray group
    (and many empty splits): the rays past a split's end must add nothing."""
    print(f"case few rays {precision} {nkeep}", flush=True)
    optic, rays = _traced(nx=64)
    v = rays.torch("vignetted")
    live = torch.nonzero(v == 0).flatten()
    v[live[nkeep:]] = 1
    point = rays.r[live[0].item()]
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    time = float(rays.t[0])
    for nx, ny in [(64, 64), (200, 136)]:
        ref = huygens_amplitudes(rays, point, prim, nx, ny, time, method='direct')
        got = huygens_amplitudes(rays, point, prim, nx, ny, time, method='gemm',
                                 precision=precision)
        assert rel_l2(got, ref) < 1e-5


def test_gemm_sheared_lattice_and_all_rays():
    """Non-diagonal primitive vectors, vignetted rays included."""
    optic, rays = _traced(nx=64, theta=0.3)
    point = rays.r[np.flatnonzero(~rays.vignetted)[5]]
    prim = np.array([[0.6e-6, 0.2e-6], [-0.1e-6, 0.5e-6]])
    time = float(rays.t[3])
    ref = huygens_amplitudes(rays, point, prim, 40, 36, time, ignoreVignetted=False,
                             method='direct')
    got = huygens_amplitudes(rays, point, prim, 40, 36, time, ignoreVignetted=False)
    assert rel_l2(got, ref) < 1e-5


def test_amplitude_matches_sumAmplitude():
    """Each pixel of the image is the reference's sumAmplitude at that point."""
    optic, rays = _traced(nx=48)
    point = rays.r[np.flatnonzero(~rays.vignetted)[0]]
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    time = float(rays.t[0])
    amp = huygens_amplitudes(rays, point, prim, 16, 16, time, method='direct')
    for (j, i) in [(0, 0), (8, 8), (3, 11), (15, 2)]:
        p = point[:2] + (i - 8) * prim[0] + (j - 8) * prim[1]
        s = rays.sumAmplitude([p[0], p[1], 0.0], time)
        assert abs(amp[j, i] - s) <= 1e-12 * abs(s) + 1e-9


def test_empty_and_all_vignetted():
    optic, rays = _traced(nx=16)
    v = rays.torch("vignetted")
    v.fill_(1)
    amp = huygens_amplitudes(rays, np.zeros(3), np.eye(2) * 1e-6, 8, 8, 0.0)
    assert np.all(amp == 0)


def test_huygensPSF_api():
    """huygensPSF end to end: the MFMA image equals the FP64 image."""
    optic = workloads.newtonian()
    kw = dict(theta_x=np.deg2rad(0.05), theta_y=0.0, wavelength=500e-9, nx=64,
              dx=1e-6, nxOut=48)
    a = bt.huygensPSF(optic, method='direct', **kw)
    b = bt.huygensPSF(optic, **kw)
    assert a.array.shape == (48, 48)
    assert rel_l2(b.array, a.array) < 2e-5
    np.testing.assert_array_equal(a.primitiveVectors, [[1e-6, 0], [0, 1e-6]])
