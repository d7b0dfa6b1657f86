"""CPU restatement of the reference's parallel-beam ray generators -- test
infrastructure only: the GPU's bt_generate_rays is checked against it and
nothing in the product imports it.

Follows batoid/rayVector.py:219-444 (asGrid), 459-643 (asPolar), 645-814
(asSpokes), 816-858 (_finish), 860-982 (fromStop) and
batoid/utils.py:483-535 (hexapolar): stop coordinates by the reference's own
numpy expressions, then the stop surface's sag, the stop -> ray frame
transform (src/batoid.cpp:8-22) and finishParallel (src/batoid.cpp:40-69)
through the C oracle.

Pins (tests/test_raygen_oracle.py): the hexapolar table against the ring
rule's closed forms and the reference's test_asPolar properties
(tests/test_RayVector.py:530-605); the lattice against the reference's
test_asGrid spacing checks (tests/test_RayVector.py:368-527); the PCG64
jump-ahead against numpy's own Generator; finishParallel against the
reference's compiled source (oracle/_ref) and the plane invariant.
"""
import ctypes

import numpy as np

from . import host as H

# ---------------------------------------------------------------------------
# stop coordinates
# ---------------------------------------------------------------------------


def hexapolar(outer, inner, nrad, naz, kfold=6):
    """(r, theta) of the hexapolar rings, outermost first (utils.py:512-531)."""
    rhos = np.linspace(outer, inner, nrad)
    nphis = []
    for rho in rhos:
        nphi = int((naz * rho / outer) // kfold) * kfold
        if nphi == 0:
            nphi = kfold
        nphis.append(nphi)
    if inner == 0.0:
        nphis[-1] = 1
    n = int(np.sum(nphis))
    rr = np.empty(n)
    th = np.empty(n)
    idx = 0
    for rho, nphi in zip(rhos, nphis):
        rr[idx:idx + nphi] = rho
        th[idx:idx + nphi] = np.linspace(0, 2 * np.pi, nphi, endpoint=False)
        idx += nphi
    if inner == 0.0:
        rr[-1] = 0.0
        th[-1] = 0.0
    return rr, th


def grid_uv(nx, ny):
    """Unit lattice coordinates of asGrid, ray i = (i // nx, i % nx)
    (rayVector.py:418-430)."""
    x_d = 1. if nx <= 2 else (nx - (2 if nx % 2 == 0 else 1)) / nx
    y_d = 1. if ny <= 2 else (ny - (2 if ny % 2 == 0 else 1)) / ny
    xx = np.fft.fftshift(np.fft.fftfreq(nx, x_d))
    yy = np.fft.fftshift(np.fft.fftfreq(ny, y_d))
    xx, yy = np.meshgrid(xx, yy)
    return xx.ravel(), yy.ravel()


def lattice_xy(lx, ly, uu, vv):
    """x = lx . (u, v), y = ly . (u, v) (rayVector.py:431-432)."""
    stack = np.stack([uu, vv])
    return np.dot(lx, stack), np.dot(ly, stack)


def polar_xy(rr, th):
    """rayVector.py:630-631."""
    return rr * np.cos(th), rr * np.sin(th)


# ---------------------------------------------------------------------------
# PCG64 stream (numpy.random.PCG64, XSL-RR 128/64), draw k by jump-ahead
# ---------------------------------------------------------------------------
PCG_MULT = (2549297995355413924 << 64) + 4865540595714422341
_MASK = (1 << 128) - 1


def _jump_tables():
    """A[b] = M^(2^b), G[b] = 1 + M + ... + M^(2^b - 1), mod 2^128."""
    a, g = PCG_MULT, 1
    A, G = [], []
    for _ in range(64):
        A.append(a)
        G.append(g)
        g = (g * (a + 1)) & _MASK
        a = (a * a) & _MASK
    return A, G


_A, _G = _jump_tables()


def pcg64_state_at(state, inc, k):
    """State after k steps, composed from the 2^b jumps of k's set bits."""
    acc_a, acc_g = 1, 0
    b = 0
    while k:
        if k & 1:
            acc_g = (_A[b] * acc_g + _G[b]) & _MASK
            acc_a = (_A[b] * acc_a) & _MASK
        k >>= 1
        b += 1
    return (acc_a * state + inc * acc_g) & _MASK


def pcg64_output(s):
    x = ((s >> 64) ^ s) & ((1 << 64) - 1)
    rot = s >> 122
    return ((x >> rot) | (x << ((64 - rot) & 63))) & ((1 << 64) - 1)


def pcg64_uniform(state, inc, k):
    """Draw k (from 0) of Generator(PCG64).random()."""
    return (pcg64_output(pcg64_state_at(state, inc, k + 1)) >> 11) * (1.0 / 9007199254740992.0)


# ---------------------------------------------------------------------------
# stop -> source plane
# ---------------------------------------------------------------------------
def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


def finish(x, y, stop_surface, stop_to_ray, dirCos, n, backDist, wavelength, flux=1.0):
    """Rays from stop coordinates (x, y): sag, stop -> ray frame, then
    finishParallel and the constant columns (rayVector.py:435-444, 816-841).
    ``stop_to_ray`` is (dr, drot) of the frame change; ``n`` the index of the
    incoming medium at ``wavelength``.  Returns oracle.host.HostRays."""
    O = H.load()
    x = np.ascontiguousarray(x, dtype=float).copy()
    y = np.ascontiguousarray(y, dtype=float).copy()
    z = np.empty_like(x)
    tab = H.host_records(stop_surface)
    O.or_sag_n(tab.ptr(), x.ctypes.data, y.ctypes.data, z.ctypes.data, len(x), 1)
    dr, drot = stop_to_ray
    O.or_apply_forward_transform_arrays(_dv(dr), _dv(drot), x.ctypes.data, y.ctypes.data,
                                        z.ctypes.data, len(x), 1)
    vv = np.array(dirCos, dtype=float, copy=True)
    vv /= n * np.sqrt(np.dot(vv, vv))
    zhat = -n * vv
    xhat = np.cross(np.array([1.0, 0.0, 0.0]), zhat)
    xhat /= np.sqrt(np.dot(xhat, xhat))
    yhat = np.cross(xhat, zhat)
    origin = zhat * backDist
    rot = np.stack([xhat, yhat, zhat]).T
    O.or_finish_parallel(_dv(origin), _dv(rot), _dv(vv), x.ctypes.data, y.ctypes.data,
                         z.ctypes.data, len(x))
    return H.HostRays(x=x, y=y, z=z, vx=vv[0], vy=vv[1], vz=vv[2], t=0.0,
                      wavelength=wavelength, flux=flux)
