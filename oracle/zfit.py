"""CPU restatement of the reference's wavefront and Zernike decompositions
-- test infrastructure only.

Follows batoid/analysis.py:308-388 (wavefront), 479-573 (zernike: galsim's
zernikeBasis, then numpy lstsq) and 576-641 (zernikeGQ: basis . (w flux) /
area), with the trace, the sphere intersection and the frame changes run by
the C oracle.

The circular Zernike basis is written here from its closed form (Noll 1976:
R_n^m(rho) = sum_s (-1)^s (n-s)! / (s! ((n+m)/2-s)! ((n-m)/2-s)!)
rho^(n-2s), normalised to unit mean square on the unit disk), independently
of batoid_amd's monomial tables; the annular basis (R_inner > 0) evaluates
batoid_amd.zernike_poly's exact Gram-Schmidt expansions, which
tests/test_host.py pins by annulus orthonormality.  Pins: the reference's
Zemax HSC Zernike coefficients and wavefront map (tests/test_zfit_oracle.py).
"""
import ctypes
from math import factorial

import numpy as np

import batoid_amd as bt
from batoid_amd.plan import PlanSpec, collect_steps
from . import host as H
from .raygen import finish, grid_uv, lattice_xy


def noll(j):
    """Noll index -> (n, |m|, 'cos' | 'sin' | None)."""
    n = 0
    while (n + 1) * (n + 2) // 2 < j:
        n += 1
    k = j - n * (n + 1) // 2 - 1
    am = n % 2 + 2 * ((k + (n + 1) % 2) // 2)
    if am == 0:
        return n, 0, None
    return n, am, ('cos' if j % 2 == 0 else 'sin')


def radial(n, m, rho):
    out = np.zeros_like(rho)
    for s in range((n - m) // 2 + 1):
        c = (-1) ** s * factorial(n - s) / (
            factorial(s) * factorial((n + m) // 2 - s) * factorial((n - m) // 2 - s))
        out += c * rho ** (n - 2 * s)
    return out


def zernike_basis(jmax, x, y, R_outer=1.0, R_inner=0.0):
    """(jmax+1, N) basis values; row 0 is zero (galsim's zernikeBasis layout,
    as analysis.py:557-560 uses it)."""
    x = np.asarray(x, float) / R_outer
    y = np.asarray(y, float) / R_outer
    out = np.zeros((jmax + 1,) + x.shape)
    if R_inner:
        from fractions import Fraction
        from batoid_amd.zernike_poly import zernike_xy_exact
        eps = Fraction(float(R_inner) / float(R_outer))
        for j in range(1, jmax + 1):
            for (i, k), v in zernike_xy_exact(j, eps).items():
                out[j] += v * x ** i * y ** k
        return out
    rho = np.hypot(x, y)
    th = np.arctan2(y, x)
    for j in range(1, jmax + 1):
        n, m, kind = noll(j)
        r = radial(n, m, rho)
        if kind is None:
            out[j] = np.sqrt(n + 1) * r
        elif kind == 'cos':
            out[j] = np.sqrt(2 * (n + 1)) * r * np.cos(m * th)
        else:
            out[j] = np.sqrt(2 * (n + 1)) * r * np.sin(m * th)
    return out


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


def _get_n(medium, w):
    tab = H.host_records(medium)
    w = np.array([w], dtype=float)
    out = np.empty(1)
    H.load().or_get_n_n(tab.ptr(), w.ctypes.data, out.ctypes.data, 1, 1)
    return out[0]


def grid_rays(optic, dirCos, wavelength, nx):
    """RayVector.asGrid(optic=..., nx=nx) with the oracle (rayVector.py:219-444)."""
    stop = optic.stopSurface
    ct = bt.CoordTransform(stop.coordSys, optic.coordSys)
    uu, vv = grid_uv(nx, nx)
    L = optic.pupilSize
    x, y = lattice_xy((L, 0.0), (0.0, L), uu, vv)
    return finish(x, y, stop.surface, (ct.dr, ct.drot), dirCos,
                  _get_n(optic.inMedium, wavelength), optic.backDist, wavelength)


def stop_xy(optic, rays):
    """Entrance-pupil positions: the rays moved to the stop frame and onto
    the stop surface (analysis.py:548-552)."""
    O = H.load()
    ep = rays.copy()
    ct = bt.CoordTransform(optic.coordSys, optic.stopSurface.coordSys)
    O.or_apply_forward_transform(_dv(ct.dr), _dv(ct.drot), ep.struct, 1)
    tab = H.host_records(optic.stopSurface.surface)
    O.or_intersect(tab.ptr(), _dv(np.zeros(3)), _dv(np.eye(3)), ep.struct, 5, H.nthreads())
    return ep["x"], ep["y"]


def wavefront(optic, theta_x, theta_y, wavelength, projection='postel', nx=32,
              sphereRadius=None, reference='chief'):
    """(waves [nx*nx], vignetted [nx*nx]) of analysis.py:308-388."""
    O = H.load()
    dirCos = bt.utils.fieldToDirCos(theta_x, theta_y, projection=projection)
    rays = grid_rays(optic, dirCos, wavelength, nx)
    if sphereRadius is None:
        sphereRadius = optic.sphereRadius
    steps = collect_steps(optic)
    H.HostPlan(PlanSpec(steps, optic.coordSys)).trace(rays)
    cs = steps[-1][0].coordSys
    cr = (nx // 2) * nx + nx // 2 if nx % 2 == 0 else (nx * nx - 1) // 2
    vig = rays["vignetted"].astype(bool)
    if reference == 'mean':
        point = np.array([np.mean(rays[k][~vig]) for k in "xyz"])
    else:
        point = np.array([rays[k][cr] for k in "xyz"])
    target = cs.shiftLocal(point + np.array([0, 0, sphereRadius]))
    ct = bt.CoordTransform(cs, target)
    O.or_apply_forward_transform(_dv(ct.dr), _dv(ct.drot), rays.struct, 1)
    tab = H.host_records(bt.Sphere(-sphereRadius))
    O.or_intersect(tab.ptr(), _dv(np.zeros(3)), _dv(np.eye(3)), rays.struct, 5, H.nthreads())
    vig = rays["vignetted"].astype(bool)
    t0 = np.mean(rays["t"][~vig]) if reference == 'mean' else rays["t"][cr]
    return (t0 - rays["t"]) / wavelength, vig


def zernike(optic, theta_x, theta_y, wavelength, projection='postel', nx=32,
            sphereRadius=None, reference='chief', jmax=22, eps=0.0):
    """analysis.py:479-573 with include_vignetted=False."""
    dirCos = bt.utils.fieldToDirCos(theta_x, theta_y, projection=projection)
    x, y = stop_xy(optic, grid_rays(optic, dirCos, wavelength, nx))
    wf, vig = wavefront(optic, theta_x, theta_y, wavelength, projection, nx, sphereRadius,
                        reference)
    w = ~vig
    basis = zernike_basis(jmax, x[w], y[w], R_outer=optic.pupilSize / 2,
                          R_inner=optic.pupilSize / 2 * eps)
    coefs, _, rank, _ = np.linalg.lstsq(basis.T, wf[w], rcond=None)
    if rank < jmax:
        raise RuntimeError(f"rank {rank} < {jmax}")
    coefs[0] = 0.0
    return coefs


def _trace_to_sphere(optic, rays, point, sphereRadius):
    """Move traced rays into the frame one radius beyond ``point`` and onto
    the reference sphere (analysis.py:624-633)."""
    O = H.load()
    cs = collect_steps(optic)[-1][0].coordSys
    target = cs.shiftLocal(point + np.array([0, 0, sphereRadius]))
    ct = bt.CoordTransform(cs, target)
    O.or_apply_forward_transform(_dv(ct.dr), _dv(ct.drot), rays.struct, 1)
    tab = H.host_records(bt.Sphere(-sphereRadius))
    O.or_intersect(tab.ptr(), _dv(np.zeros(3)), _dv(np.eye(3)), rays.struct, 5, H.nthreads())


def zernike_gq(optic, theta_x, theta_y, wavelength, projection='postel', rings=6,
               spokes=None, sphereRadius=None, reference='chief', jmax=22, eps=0.0):
    """analysis.py:576-641: Gauss-quadrature rays (rayVector.py:785-794),
    traced, onto the reference sphere, projected on the basis."""
    dirCos = bt.utils.fieldToDirCos(theta_x, theta_y, projection=projection)
    outer = optic.pupilSize / 2
    inner = eps * outer
    if spokes is None:
        spokes = 2 * rings + 1
    Li, w = np.polynomial.legendre.leggauss(rings)
    e = inner / outer
    area = np.pi * (1 - e**2)
    rr = np.sqrt(e**2 + (1 + Li) * (1 - e**2) / 2) * outer
    flux = w * area / (2 * spokes)
    sp = np.linspace(0, 2 * np.pi, spokes, endpoint=False)
    rr, sp = np.meshgrid(rr, sp)
    flux = np.broadcast_to(flux, rr.shape).ravel()
    x, y = rr.ravel() * np.cos(sp.ravel()), rr.ravel() * np.sin(sp.ravel())
    stop = optic.stopSurface
    ct = bt.CoordTransform(stop.coordSys, optic.coordSys)
    nidx = _get_n(optic.inMedium, wavelength)
    rays = finish(x, y, stop.surface, (ct.dr, ct.drot), dirCos, nidx, optic.backDist,
                  wavelength, flux)
    ex, ey = stop_xy(optic, rays)
    if sphereRadius is None:
        sphereRadius = optic.sphereRadius
    plan = H.HostPlan(PlanSpec(collect_steps(optic), optic.coordSys))
    plan.trace(rays)
    assert not rays["failed"].any()
    vig = rays["vignetted"].astype(bool)
    if reference == 'mean':
        point = np.array([np.mean(rays[k][~vig]) for k in "xyz"])
    else:
        chief = finish([0.0], [0.0], stop.surface, (ct.dr, ct.drot), dirCos, nidx,
                       optic.backDist, wavelength)
        plan.trace(chief)
        point = np.array([chief[k][0] for k in "xyz"])
    _trace_to_sphere(optic, rays, point, sphereRadius)
    if reference == 'mean':
        t0 = np.mean(rays["t"][~rays["vignetted"].astype(bool)])
    else:
        _trace_to_sphere(optic, chief, point, sphereRadius)
        t0 = chief["t"][0]
    basis = zernike_basis(jmax, ex, ey, R_outer=outer, R_inner=inner)
    return np.dot(basis, (t0 - rays["t"]) / wavelength * rays["flux"]) / (np.pi * (1. - eps**2))
