"""Huygens PSF, wavefront (OPD), its Zernike decomposition and the
plate-scale Jacobian (batoid/analysis.py:10-56, 181-388, 479-641) on the
GPU.

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
        k = _chief_index(nx)             # one ray's position read back, not the grid's
        return torch.stack([rays.torch(c)[k] for c in ("x", "y", "z")]).cpu().numpy()
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


def _sphere_rays(optic, theta_x, theta_y, wavelength, projection, nx, sphereRadius,
                 reference):
    """The wavefront grid traced through the optic and onto the reference
    sphere (analysis.py:342-373); returns the rays (device) and t0."""
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
    if reference == 'mean':
        t = rays.t
        t0 = np.mean(t[np.where(1 - rays.vignetted)[0]])
    else:
        t0 = float(rays.torch("t")[_chief_index(nx)].item())
    return rays, t0


def wavefront(optic, theta_x, theta_y, wavelength, projection='postel', nx=32,
              sphereRadius=None, reference='chief'):
    """Wavefront in waves against a reference sphere centred on the image
    point (analysis.py:308-388)."""
    rays, t0 = _sphere_rays(optic, theta_x, theta_y, wavelength, projection, nx,
                            sphereRadius, reference)
    arr = np.ma.masked_array((t0 - rays.t) / wavelength, mask=rays.vignetted).reshape(nx, nx)
    d = nx - 2 if (nx % 2) == 0 else nx - 1
    primitiveU = np.vstack([[optic.pupilSize / d, 0], [0, optic.pupilSize / d]])
    return Lattice(arr, primitiveU)


def zernike_table(jmax, eps=0.0):
    """Monomial coefficients of the Noll Zernikes Z_1 .. Z_jmax (unit outer
    radius, obscuration ``eps``) as a [jmax][M] table, column
    m = d(d+1)/2 + q for x^(d-q) y^q; returns (table, radial order)."""
    from fractions import Fraction
    from .zernike_poly import noll_to_nm, zernike_xy_exact
    nmax = max(noll_to_nm(j)[0] for j in range(1, jmax + 1))
    C = np.zeros((jmax, (nmax + 1) * (nmax + 2) // 2))
    e = Fraction(float(eps)) if eps else Fraction(0)
    for j in range(1, jmax + 1):
        for (i, k), v in zernike_xy_exact(j, e).items():
            d = i + k
            C[j - 1, d * (d + 1) // 2 + k] = v
    return C, nmax


_COEF_CACHE = {}     # (jmax, eps, device) -> (monomial table on the device, order)


def zernike_gram(x, y, t, t0, wavelength, R_outer, R_inner=0.0, jmax=22, flux=None,
                 mask=None):
    """Gram matrix of [Z_1 .. Z_jmax, w] over the rays not masked, with
    w = (t0 - t) / wavelength (times ``flux`` if given) and the Zernikes
    evaluated at (x, y) / R_outer -- one bt_zernike_gram pass over device
    columns.  Returns the (jmax+1, jmax+1) symmetric matrix (host)."""
    if not 1 <= jmax <= 78:
        raise ValueError("jmax must be in [1, 78]")
    eps = R_inner / R_outer if R_inner else 0.0
    lib = _lib.load()
    dev = _device.device()
    key = (jmax, eps, dev.index)
    if key not in _COEF_CACHE:
        C, nmax = zernike_table(jmax, eps)
        _COEF_CACHE[key] = (_device.f64(C), nmax)
    coef, nmax = _COEF_CACHE[key]
    ws = torch.empty(int(lib.bt_zernike_workspace_bytes(jmax)), dtype=torch.uint8, device=dev)
    nent = (jmax + 1) * (jmax + 2) // 2
    out = torch.empty(nent, dtype=torch.float64, device=dev)
    n = t.numel()
    _lib.call("bt_zernike_gram", _device.ptr(x), _device.ptr(y), _device.ptr(t),
              _device.ptr(flux), _device.ptr(mask), n, float(t0), float(wavelength),
              float(R_outer), _device.ptr(coef), jmax, nmax, _device.ptr(out),
              _device.ptr(ws), ws.numel(), _device.stream())
    packed = _device.to_numpy(out)
    G = np.zeros((jmax + 1, jmax + 1))
    G[np.triu_indices(jmax + 1)] = packed
    return G + np.triu(G, 1).T


def zernike(optic, theta_x, theta_y, wavelength, projection='postel', nx=32,
            sphereRadius=None, reference='chief', jmax=22, eps=0.0,
            include_vignetted=False):
    """Least-squares Zernike decomposition of the wavefront, in waves
    (analysis.py:479-573).  The normal equations are accumulated on the GPU
    (bt_zernike_gram) and the jmax x jmax system is solved on the host;
    index 0 of the result is 0 (Noll indices start at 1)."""
    dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
    rays = RayVector.asGrid(optic=optic, wavelength=wavelength, nx=nx, dirCos=dirCos)
    epRays = rays.toCoordSys(optic.stopSurface.coordSys)
    optic.stopSurface.surface.intersect(epRays)
    srays, t0 = _sphere_rays(optic, theta_x, theta_y, wavelength, projection, nx,
                             sphereRadius, reference)
    if include_vignetted:
        d = nx - 2 if (nx % 2) == 0 else nx - 1
        dx = optic.pupilSize / d
        k = np.arange(nx) - nx // 2
        cx, cy = np.meshgrid(k * dx, k * dx)
        mask = _device.u8((np.hypot(cx, cy) > optic.pupilSize / 2).ravel())
    else:
        mask = srays.torch("vignetted")
    G = zernike_gram(epRays.torch("x"), epRays.torch("y"), srays.torch("t"), t0, wavelength,
                     optic.pupilSize / 2, optic.pupilSize / 2 * eps, jmax, mask=mask)
    return solve_normal_equations(G, jmax)


def solve_normal_equations(G, jmax):
    """Least-squares Zernike coefficients from the Gram matrix G of
    [Z_1 .. Z_jmax, w] (zernike_gram), with the reference's rank check
    (analysis.py:566-570: RuntimeError when the basis has rank < jmax).

    The basis B (rows = unmasked rays) is not formed; its rank is judged on
    A = B^T B.  Z_1 = 1, so A[0, 0] is the row count N: N < jmax is rank
    deficient outright.  Otherwise an eigenvalue of A counts when it exceeds
    lambda_max * max(N, jmax + 1) * eps -- a singular value of B above
    s_max * sqrt(max(N, jmax + 1) * eps), since A's eigenvalues carry absolute
    errors of order eps * lambda_max and smaller singular values cannot be
    resolved from A.  (numpy's lstsq on B resolves down to
    s_max * max(N, jmax + 1) * eps: a basis whose smallest singular value lies
    between the two, condition number above ~1e6, is accepted by the
    reference and rejected here.)"""
    A, b = G[:jmax, :jmax], G[:jmax, jmax]
    nrows = int(round(G[0, 0]))            # Z_1 = 1: the count of unmasked rays
    if nrows < jmax:
        raise RuntimeError(
            f"Rank {nrows} matrix found when fitting {jmax} Zernike coefficients.  ")
    lam = np.linalg.eigvalsh(A)
    tol = lam.max() * max(nrows, jmax + 1) * np.finfo(float).eps
    rank = int((lam > tol).sum()) if lam.max() > 0 else 0
    if rank < jmax:
        raise RuntimeError(
            f"Rank {rank} matrix found when fitting {jmax} Zernike coefficients.  ")
    coefs = np.zeros(jmax + 1)
    coefs[1:] = np.linalg.solve(A, b)
    return coefs


def zernikeGQ(optic, theta_x, theta_y, wavelength, projection='postel', rings=6,
              spokes=None, sphereRadius=None, reference='chief', jmax=22, eps=0.0):
    """Zernike decomposition by Gaussian quadrature over the (annular) pupil,
    in waves (analysis.py:576-641): flux-weighted projections of the
    wavefront onto each Zernike, one bt_zernike_gram pass."""
    from .surface import Sphere
    dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
    inner = eps * optic.pupilSize / 2
    rays = RayVector.asSpokes(optic=optic, wavelength=wavelength, inner=inner, dirCos=dirCos,
                              rings=rings, spokes=spokes, spacing='GQ')
    epRays = rays.copy().toCoordSys(optic.stopSurface.coordSys)
    optic.stopSurface.surface.intersect(epRays)
    if sphereRadius is None:
        sphereRadius = optic.sphereRadius
    optic.trace(rays)
    if bool(rays.torch("failed").any()):
        raise ValueError("Cannot compute zernike with Gaussian Quadrature with failed rays.")
    if reference == 'mean':
        w = np.where(1 - rays.vignetted)[0]
        point = np.mean(rays.r[w], axis=0)
    elif reference == 'chief':
        chiefRay = RayVector.fromStop(0.0, 0.0, backDist=optic.backDist, wavelength=wavelength,
                                      dirCos=dirCos, medium=optic.inMedium,
                                      stopSurface=optic.stopSurface)
        optic.trace(chiefRay)
        point = chiefRay.r[0]
    else:
        raise ValueError(f"unknown reference {reference!r}")
    target = rays.coordSys.shiftLocal(point + np.array([0, 0, sphereRadius]))
    rays.toCoordSys(target)
    sphere = Sphere(-sphereRadius)
    sphere.intersect(rays)
    if reference == 'mean':
        w = np.where(1 - rays.vignetted)[0]
        t0 = np.mean(rays.t[w])
    else:
        chiefRay.toCoordSys(target)
        sphere.intersect(chiefRay)
        t0 = float(chiefRay.t[0])
    G = zernike_gram(epRays.torch("x"), epRays.torch("y"), rays.torch("t"), t0, wavelength,
                     optic.pupilSize / 2, inner, jmax, flux=rays.torch("flux"))
    area = np.pi * (1. - eps**2)
    out = np.zeros(jmax + 1)
    out[1:] = G[:jmax, jmax] / area
    return out
