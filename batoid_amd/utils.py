"""Field-angle projections, hexapolar sampling and the bilinear fit
(batoid/utils.py).  Host-side numpy, as in the reference."""
import numpy as np


def normalized(args):
    args = np.atleast_1d(args)
    return args / np.linalg.norm(args)


def bilinear_fit(ux, uy, kx, ky):
    """Least-squares fit k = a + b ux + c uy for both components
    (utils.py:9-18); rows of the result are (a, b, c)."""
    a = np.empty((len(ux), 3), dtype=float)
    a[:, 0] = 1
    a[:, 1] = ux
    a[:, 2] = uy
    b = np.empty((len(ux), 2), dtype=float)
    b[:, 0] = kx
    b[:, 1] = ky
    x, _, _, _ = np.linalg.lstsq(a, b, rcond=-1)
    return x


def _zero_guard(rho, alpha, beta, gamma, g0):
    w = (rho == 0.0)
    if np.ndim(w) == 0:
        return (0.0, 0.0, g0) if w else (alpha, beta, gamma)
    if np.any(w):
        alpha = np.array(alpha)
        beta = np.array(beta)
        gamma = np.array(gamma)
        alpha[w] = 0.0
        beta[w] = 0.0
        gamma[w] = g0
    return alpha, beta, gamma


def postelToDirCos(u, v):
    """Postel (azimuthal equidistant) tangent-plane coordinates -> direction
    cosines; (0, 0) is a ray travelling along -z (utils.py:72-109)."""
    rho = np.sqrt(u * u + v * v)
    if np.ndim(rho) == 0 and rho == 0.0:
        return 0.0, 0.0, -1.0
    srho = np.sin(rho)
    with np.errstate(invalid='ignore', divide='ignore'):
        alpha = u / rho * srho
        beta = v / rho * srho
    gamma = -np.cos(rho)
    return _zero_guard(rho, alpha, beta, gamma, -1.0)


def dirCosToPostel(alpha, beta, gamma):
    w = (gamma == -1)
    if np.ndim(w) == 0 and w:
        return 0.0, 0.0
    rho = np.arccos(-gamma)
    srho = np.sin(rho)
    with np.errstate(invalid='ignore', divide='ignore'):
        u = alpha * rho / srho
        v = beta * rho / srho
    if np.ndim(w) and np.any(w):
        u[w] = 0.0
        v[w] = 0.0
    return u, v


def zemaxToDirCos(u, v):
    """Zemax field angles -> direction cosines (utils.py:149-175)."""
    tanu = np.tan(u)
    tanv = np.tan(v)
    norm = np.sqrt(1 + tanu * tanu + tanv * tanv)
    return tanu / norm, tanv / norm, -1 / norm


def dirCosToZemax(alpha, beta, gamma):
    return np.arctan(-alpha / gamma), np.arctan(-beta / gamma)


def gnomonicToDirCos(u, v):
    """Gnomonic tangent-plane coordinates -> direction cosines (utils.py:21-40)."""
    gamma = 1 / np.sqrt(1.0 + u * u + v * v)
    return u * gamma, v * gamma, -gamma


def dirCosToGnomonic(alpha, beta, gamma):
    return -alpha / gamma, -beta / gamma


def stereographicToDirCos(u, v):
    """Stereographic tangent-plane coordinates -> direction cosines."""
    rho = np.sqrt(u * u + v * v)
    if np.ndim(rho) == 0 and rho == 0.0:
        return 0.0, 0.0, -1.0
    theta = 2 * np.arctan(rho / 2)
    stheta = np.sin(theta)
    with np.errstate(invalid='ignore', divide='ignore'):
        alpha = u / rho * stheta
        beta = v / rho * stheta
    gamma = -np.cos(theta)
    return _zero_guard(rho, alpha, beta, gamma, -1.0)


def dirCosToStereographic(alpha, beta, gamma):
    w = (gamma == -1)
    if np.ndim(w) == 0 and w:
        return 0.0, 0.0
    theta = np.arccos(-gamma)
    rho = 2 * np.tan(theta / 2)
    stheta = np.sin(theta)
    with np.errstate(invalid='ignore', divide='ignore'):
        u = alpha * rho / stheta
        v = beta * rho / stheta
    if np.ndim(w) and np.any(w):
        u[w] = 0.0
        v[w] = 0.0
    return u, v


def orthographicToDirCos(u, v):
    """Orthographic tangent-plane coordinates -> direction cosines."""
    rho = np.sqrt(u * u + v * v)
    theta = np.arcsin(rho)
    gamma = np.cos(theta)
    return u, v, -gamma


def dirCosToOrthographic(alpha, beta, gamma):
    return alpha, beta


def lambertToDirCos(u, v):
    """Lambert azimuthal equal-area coordinates -> direction cosines.  At the
    origin the reference returns gamma = -1 for scalars and +1 inside
    arrays (utils.py:340-374); both are kept."""
    rhosqr = u * u + v * v
    if np.ndim(rhosqr) == 0 and rhosqr == 0.0:
        return 0.0, 0.0, -1.0
    rho = np.sqrt(rhosqr)
    gamma = (2 - rhosqr) / 2
    r = np.sqrt(1 - gamma * gamma)
    with np.errstate(invalid='ignore', divide='ignore'):
        alpha = u * r / rho
        beta = v * r / rho
    alpha, beta, gamma = _zero_guard(rhosqr, alpha, beta, gamma, 1.0)
    return alpha, beta, -gamma


def dirCosToLambert(alpha, beta, gamma):
    w = (gamma == -1)
    if np.ndim(w) == 0 and w:
        return 0.0, 0.0
    rho = np.sqrt(2 + 2 * gamma)
    norm = np.sqrt(1 - gamma * gamma)
    with np.errstate(invalid='ignore', divide='ignore'):
        u = alpha * rho / norm
        v = beta * rho / norm
    if np.ndim(w) and np.any(w):
        u[w] = 0.0
        v[w] = 0.0
    return u, v


_TO = {'postel': postelToDirCos, 'zemax': zemaxToDirCos,
       'gnomonic': gnomonicToDirCos, 'stereographic': stereographicToDirCos,
       'lambert': lambertToDirCos, 'orthographic': orthographicToDirCos}
_FROM = {'postel': dirCosToPostel, 'zemax': dirCosToZemax,
         'gnomonic': dirCosToGnomonic, 'stereographic': dirCosToStereographic,
         'lambert': dirCosToLambert, 'orthographic': dirCosToOrthographic}


def fieldToDirCos(u, v, projection='postel'):
    """Field angle (radians) -> direction cosines (utils.py:409-443)."""
    if projection not in _TO:
        raise ValueError("Bad projection: {}".format(projection))
    return _TO[projection](u, v)


def dirCosToField(alpha, beta, gamma, projection='postel'):
    if projection not in _FROM:
        raise ValueError("Bad projection: {}".format(projection))
    return _FROM[projection](alpha, beta, gamma)


def hexapolar(outer=1.0, inner=0.0, nrad=5, naz=30, kfold=6, rth=False):
    """Hexapolar grid: ``nrad`` rings from outer to inner, each with a
    multiple of ``kfold`` azimuths in proportion to its radius
    (utils.py:483-535)."""
    rhos = np.linspace(outer, inner, nrad)
    nphis = []
    for rho in rhos:
        nphi = int((naz * rho / outer) // kfold) * kfold
        if nphi == 0:
            nphi = kfold
        nphis.append(nphi)
    if inner == 0.0:
        nphis[-1] = 1
    n = np.sum(nphis)
    th = np.empty(n)
    rr = np.empty(n)
    idx = 0
    for rho, nphi in zip(rhos, nphis):
        rr[idx:idx + nphi] = rho
        th[idx:idx + nphi] = np.linspace(0, 2 * np.pi, nphi, endpoint=False)
        idx += nphi
    if inner == 0.0:
        rr[-1] = 0.0
        th[-1] = 0.0
    if rth:
        return rr, th
    return rr * np.cos(th), rr * np.sin(th)
