"""Annular Zernike polynomials as 2-D monomial coefficient arrays.

The reference's Zernike surface (batoid/surface.py:463-483) takes from
galsim the monomial arrays ``c[i][k]`` (coefficient of x^i y^k) of a series
sum_j a_j Z_j(x/R_outer, y/R_outer; eps = R_inner/R_outer) and of its x and
y gradients, and evaluates them with 2-D Horner on the device.  galsim is
not a dependency here, so the conversion is written from the mathematics:

* Noll (1976) indexing: j -> (n, m); even j carries cos(m theta), odd j
  sin(m theta) for m != 0.
* Radial polynomials: for eps = 0 the classical R_n^m; for eps > 0
  Mahajan's annular polynomials, i.e. the Gram-Schmidt orthogonalisation of
  rho^m, rho^(m+2), ..., rho^n on [eps, 1] with weight rho, positive leading
  coefficient, normalised so every Z_j has unit mean square on the annulus.
  The orthogonalisation runs in exact rational arithmetic (eps is converted
  exactly), with one final floating-point scale.
* rho^k cos(m theta) = (x^2+y^2)^((k-m)/2) Re((x+iy)^m), likewise Im for
  sin, expanded binomially.

Parity with galsim is unpinned at the bit level (galsim is unavailable);
tests pin these arrays with closed forms (Noll Z1..Z11, Mahajan's annular
Z4) and annulus orthonormality by quadrature.
"""
from fractions import Fraction
from functools import lru_cache
from math import comb, sqrt

import numpy as np


def noll_to_nm(j):
    """Noll index j >= 1 -> (n, m), m < 0 for the sin terms."""
    if j < 1:
        raise ValueError("Noll indices start at 1")
    n = 0
    while (n + 1) * (n + 2) // 2 < j:
        n += 1
    k = j - n * (n + 1) // 2 - 1
    am = 2 * ((k + 1) // 2) if n % 2 == 0 else 2 * (k // 2) + 1
    if am == 0:
        return n, 0
    return (n, am) if j % 2 == 0 else (n, -am)


def _moment(p, eps):
    # integral_eps^1 rho^p rho d(rho), exact
    return (1 - eps ** (p + 2)) / (p + 2)


@lru_cache(maxsize=None)
def _radial_exact(n, m, eps):
    """Monic orthogonal radial polynomial of degree n (same parity as m) on
    [eps, 1], as a dict {power: Fraction}, plus its squared norm."""
    powers = list(range(m, n + 1, 2))
    basis = []           # orthogonal polynomials so far (dicts)
    norms = []

    def inner(a, b):
        return sum(ca * cb * _moment(pa + pb, eps)
                   for pa, ca in a.items() for pb, cb in b.items())

    for p in powers:
        v = {p: Fraction(1)}
        for b, nb in zip(basis, norms):
            proj = inner(v, b) / nb
            for q, c in b.items():
                v[q] = v.get(q, Fraction(0)) - proj * c
        basis.append(v)
        norms.append(inner(v, v))
    return basis[-1], norms[-1]


@lru_cache(maxsize=None)
def zernike_xy_exact(j, eps):
    """Monomial coefficients {(i, k): float} of the unit-radius annular
    Zernike Z_j with obscuration ``eps`` (a Fraction)."""
    n, m = noll_to_nm(j)
    am = abs(m)
    poly, norm2 = _radial_exact(n, am, eps)
    # target: integral R^2 rho drho = (1 - eps^2) / (2 (n+1)), with the
    # azimuthal normalisation sqrt(n+1) (m = 0) or sqrt(2(n+1)) (m != 0)
    target = (1 - eps * eps) / (2 * (n + 1))
    scale = sqrt(float(target / norm2))
    scale *= sqrt(n + 1) if am == 0 else sqrt(2 * (n + 1))
    # angular part Re/Im (x + i y)^am as {(i, k): int}
    ang = {}
    for p in range(am + 1):
        if m >= 0 and p % 2 == 0:
            ang[(am - p, p)] = comb(am, p) * (-1) ** (p // 2)
        elif m < 0 and p % 2 == 1:
            ang[(am - p, p)] = comb(am, p) * (-1) ** ((p - 1) // 2)
    out = {}
    for power, c in poly.items():
        q = (power - am) // 2
        for r in range(q + 1):
            cr = comb(q, r)
            for (ai, ak), ca in ang.items():
                key = (ai + 2 * r, ak + 2 * (q - r))
                out[key] = out.get(key, Fraction(0)) + c * cr * ca
    return {key: float(v) * scale for key, v in out.items() if v != 0}


def zernike_coef_arrays(coef, R_outer=1.0, R_inner=0.0):
    """(c, gx, gy) monomial arrays of sum_j coef[j] Z_j(x/R_outer, y/R_outer)
    for j >= 1 (coef[0] is ignored, as in the Noll convention).

    c has shape (n+1, n+1) with n the largest radial order present; gx, gy
    (n, n) hold d/dx and d/dy.  For n = 0 the gradient arrays are a single
    zero, which 2-D Horner evaluates to 0."""
    coef = np.asarray(coef, dtype=float)
    jmax = len(coef) - 1
    nmax = max([noll_to_nm(j)[0] for j in range(1, jmax + 1)] + [0])
    eps = Fraction(float(R_inner) / float(R_outer)) if R_inner else Fraction(0)
    c = np.zeros((nmax + 1, nmax + 1))
    for j in range(1, jmax + 1):
        a = coef[j]
        if a == 0.0:
            continue
        for (i, k), v in zernike_xy_exact(j, eps).items():
            c[i, k] += a * v
    powers = np.add.outer(np.arange(nmax + 1), np.arange(nmax + 1))
    c = c / float(R_outer) ** powers
    if nmax == 0:
        return c, np.zeros((1, 1)), np.zeros((1, 1))
    gx = np.zeros((nmax, nmax))
    gy = np.zeros((nmax, nmax))
    for i in range(nmax):
        for k in range(nmax):
            gx[i, k] = (i + 1) * c[i + 1, k]
            gy[i, k] = (k + 1) * c[i, k + 1]
    return c, gx, gy
