"""Huygens PSF, wavefront (OPD) and plate-scale Jacobian
(batoid/analysis.py:10-56, 181-388) on the GPU.

``huygensPSF`` keeps the reference's signature and lattice conventions.  The
reference sums exp(i phi) over rays for every pixel in a Python loop
(analysis.py:302-303); here the whole image is one GPU call:

* ``method='gemm'`` (default): the sum factorised as a complex GEMM on MFMA
  (``precision='fp32'`` or ``'bf16x3'``), phases reduced mod 2 pi in FP64;
* ``method='direct'``: the reference's FP64 per-(pixel, ray) arithmetic.
"""
import numpy as np
import torch

from . import _lib, _device
from .lattice import Lattice
from .rayVector import RayVector
from .utils import bilinear_fit, fieldToDirCos


def _reciprocalLatticeVectors(a1, a2, N):
    """Reciprocal lattice of (a1, a2) scaled by N (analysis.py:10-14); the
    reference builds b1 from (a2[1], a2[0]) without a sign change, kept."""
    norm = 2 * np.pi / (a1[0] * a2[1] - a1[1] * a2[0]) / N
    b1 = norm * np.array([a2[1], a2[0]])
    b2 = norm * np.array([a1[1], a1[0]])
    return b1, b2


def dkdu(optic, theta_x, theta_y, wavelength, nrad=6, naz=36, projection='postel'):
    """d(kx, ky)/d(pupil u) of outgoing rays by a bilinear fit
    (analysis.py:17-56)."""
    rays = RayVector.asPolar(optic=optic, theta_x=theta_x, theta_y=theta_y,
                             wavelength=wavelength, projection=projection,
                             nrad=nrad, naz=naz)
    ux = np.array(rays.x)
    uy = np.array(rays.y)
    optic.trace(rays)
    w = ~rays.vignetted
    k = rays.k
    soln = bilinear_fit(ux[w], uy[w], k[w, 0], k[w, 1])
    return soln[1:]


def _chief_index(nx):
    return (nx // 2) * nx + nx // 2 if (nx % 2) == 0 else (nx * nx - 1) // 2


def _reference_point(rays, nx, reference):
    if reference == 'mean':
        w = np.where(1 - rays.vignetted)[0]
        return np.mean(rays.r[w], axis=0)
    if reference == 'chief':
        return rays.r[_chief_index(nx)]
    raise ValueError(f"unknown reference {reference!r}")


_PRECISION = {'fp32': _lib.PSF_FP32, 'bf16x3': _lib.PSF_BF16X3}


def huygens_amplitudes(rays, point, primitiveX, nx, ny, time, ignoreVignetted=True,
                       method='gemm', precision='fp32'):
    """Complex amplitudes [ny, nx] of the Huygens sum over ``rays`` on the
    lattice ``primitiveX`` centred on ``point`` (z = 0 plane, time ``time``).
    Row j, column i sits at point + (i - nx//2) a1 + (j - ny//2) a2."""
    lib = _lib.load()
    p0 = _lib.dvec([point[0], point[1]])
    a1 = _lib.dvec(primitiveX[0])
    a2 = _lib.dvec(primitiveX[1])
    nbytes = int(lib.bt_huygens_workspace_bytes(len(rays), nx, ny))
    ws = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=_device.device())
    out = _device.empty_f64(2 * nx * ny)
    if method == 'direct':
        _lib.call("bt_huygens_direct", rays._rays, p0, a1, a2, nx, ny, float(time),
                  int(bool(ignoreVignetted)), _device.ptr(out), _device.ptr(ws), nbytes,
                  _device.stream())
    elif method == 'gemm':
        _lib.call("bt_huygens_gemm", rays._rays, p0, a1, a2, nx, ny, float(time),
                  int(bool(ignoreVignetted)), _PRECISION[precision], _device.ptr(out),
                  _device.ptr(ws), nbytes, _device.stream())
    else:
        raise ValueError(f"unknown method {method!r}")
    return _device.to_numpy(out).view(np.complex128).reshape(ny, nx)


def huygensPSF(optic, theta_x, theta_y, wavelength, projection='postel', nx=None,
               dx=None, dy=None, nxOut=None, reference='chief', method='gemm',
               precision='fp32'):
    """PSF by the Huygens construction (analysis.py:181-305): |sum_u
    exp(i phi(u))|^2 on a lattice around the chief ray (or the mean
    unvignetted ray)."""
    from numbers import Real
    if dx is None:
        d = nx - 2 if (nx % 2) == 0 else nx - 1
        primitiveU = np.array([[optic.pupilSize / d, 0], [0, optic.pupilSize / d]])
        primitiveK = dkdu(optic, theta_x, theta_y, wavelength,
                          projection=projection).dot(primitiveU)
        pad_factor = 2
        primitiveX = np.vstack(_reciprocalLatticeVectors(
            primitiveK[0], primitiveK[1], pad_factor * nx))
    elif isinstance(dx, Real):
        if dy is None:
            dy = dx
        primitiveX = np.vstack([[dx, 0], [0, dy]])
        pad_factor = 1
    else:
        primitiveX = np.vstack([dx, dy])
        pad_factor = 1
    if nxOut is None:
        nxOut = nx
    dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
    rays = RayVector.asGrid(optic=optic, wavelength=wavelength, dirCos=dirCos, nx=nx)
    optic.trace(rays)
    point = _reference_point(rays, nx, reference)
    time = float(rays.torch("t")[0].item())
    n = nxOut * pad_factor
    amp = huygens_amplitudes(rays, point, primitiveX, n, n, time, True, method, precision)
    return Lattice(np.abs(amp) ** 2, primitiveX)


def wavefront(optic, theta_x, theta_y, wavelength, projection='postel', nx=32,
              sphereRadius=None, reference='chief'):
    """Wavefront in waves against a reference sphere centred on the image
    point (analysis.py:308-388)."""
    from .surface import Sphere
    dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
    rays = RayVector.asGrid(optic=optic, wavelength=wavelength, nx=nx, dirCos=dirCos)
    if sphereRadius is None:
        sphereRadius = optic.sphereRadius
    optic.trace(rays)
    point = _reference_point(rays, nx, reference)
    target = rays.coordSys.shiftLocal(point + np.array([0, 0, sphereRadius]))
    rays.toCoordSys(target)
    Sphere(-sphereRadius).intersect(rays)
    t = rays.t
    vig = rays.vignetted
    if reference == 'mean':
        t0 = np.mean(t[np.where(1 - vig)[0]])
    else:
        t0 = t[_chief_index(nx)]
    arr = np.ma.masked_array((t0 - t) / wavelength, mask=vig).reshape(nx, nx)
    d = nx - 2 if (nx % 2) == 0 else nx - 1
    primitiveU = np.vstack([[optic.pupilSize / d, 0], [0, optic.pupilSize / d]])
    return Lattice(arr, primitiveU)
