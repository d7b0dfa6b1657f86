"""Pin the wavefront / Zernike oracle (oracle/zfit.py): closed forms and
orthonormality of its basis, agreement with the monomial tables the GPU
uses, and the reference's Zemax HSC golden data -- Zernike coefficients
(tests/test_zemax.py:188-220) and the wavefront map (154-185).  CPU only."""
import os

import numpy as np
import pytest

import batoid_amd as bt
from batoid_amd.analysis import zernike_table
from oracle import zfit as Z

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _disk_quadrature(nr=24, nth=64, eps=0.0):
    """Gauss-Legendre in r^2 x uniform azimuth: exact for polynomials of the
    degrees used here on the (annular) unit disk; weights sum to 1."""
    t, w = np.polynomial.legendre.leggauss(nr)
    r2 = eps**2 + (t + 1) / 2 * (1 - eps**2)
    th = np.arange(nth) * 2 * np.pi / nth
    R, TH = np.meshgrid(np.sqrt(r2), th)
    W = np.broadcast_to(w / 2 / nth, R.shape)
    return (R * np.cos(TH)).ravel(), (R * np.sin(TH)).ravel(), W.ravel()


def test_basis_closed_forms():
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-0.7, 0.7, (2, 500))
    r2 = x * x + y * y
    B = Z.zernike_basis(11, x, y)
    np.testing.assert_array_equal(B[0], 0.0)
    np.testing.assert_allclose(B[1], 1.0)
    np.testing.assert_allclose(B[2], 2 * x, atol=1e-14)
    np.testing.assert_allclose(B[3], 2 * y, atol=1e-14)
    np.testing.assert_allclose(B[4], np.sqrt(3) * (2 * r2 - 1), atol=1e-14)
    np.testing.assert_allclose(B[5], 2 * np.sqrt(6) * x * y, atol=1e-14)
    np.testing.assert_allclose(B[6], np.sqrt(6) * (x * x - y * y), atol=1e-14)
    np.testing.assert_allclose(B[11], np.sqrt(5) * (6 * r2 * r2 - 6 * r2 + 1), atol=1e-13)


def test_basis_orthonormal_on_disk():
    x, y, w = _disk_quadrature()
    B = Z.zernike_basis(45, x, y)[1:]
    np.testing.assert_allclose((B * w) @ B.T, np.eye(45), atol=1e-12)


def test_monomial_tables_match_closed_form():
    """The GPU's monomial tables (circular) against the oracle's
    independent closed form, j <= 78 (radial order 11)."""
    rng = np.random.default_rng(2)
    x, y = rng.uniform(-0.7, 0.7, (2, 400))
    C, nmax = zernike_table(78)
    mono = np.array([x ** (d - q) * y ** q for d in range(nmax + 1) for q in range(d + 1)])
    np.testing.assert_allclose(C @ mono, Z.zernike_basis(78, x, y)[1:], rtol=0, atol=1e-11)
    # and annular tables are orthonormal on their annulus
    xa, ya, wa = _disk_quadrature(eps=0.4)
    Ca, na = zernike_table(37, 0.4)
    mono = np.array([xa ** (d - q) * ya ** q for d in range(na + 1) for q in range(d + 1)])
    Ba = Ca @ mono
    np.testing.assert_allclose((Ba * wa) @ Ba.T, np.eye(37), atol=1e-11)


def _hsc():
    return bt.Optic.fromYaml(os.path.join(GOLD, "optics", "HSC_no_obsc.yaml"))


def test_oracle_zernike_vs_zemax_hsc():
    """tests/test_zemax.py:188-220: HSC, 0.75 deg, 750 nm, nx=256, j<=37."""
    golden = np.load(os.path.join(GOLD, "hsc_zemax.npz"))["zernike"]
    coefs = Z.zernike(_hsc(), 0.0, np.deg2rad(0.75), 750e-9, jmax=37, nx=256,
                      projection='zemax', reference='chief')
    np.testing.assert_allclose(coefs[4:], golden[4:], rtol=0, atol=1e-2)
    np.testing.assert_allclose(coefs[11:], golden[11:], rtol=0, atol=3e-3)


def test_oracle_wavefront_vs_zemax_hsc():
    """tests/test_zemax.py:154-185: the 512 x 512 wavefront map, compared
    through Zernike fits j = 4 .. 37 on Zemax's mask, 0.01 waves."""
    Zwf = np.load(os.path.join(GOLD, "hsc_zemax.npz"))["wavefront"][::-1]
    nx = 512
    wf, vig = Z.wavefront(_hsc(), 0.0, np.deg2rad(0.75), 750e-9, nx=nx)
    wf = wf.reshape(nx, nx)
    x = np.linspace(-1, 1, nx, endpoint=False)
    x, y = np.meshgrid(x, x)
    w = Zwf != 0
    basis = Z.zernike_basis(37, x[w], y[w])
    zc = np.linalg.lstsq(basis.T, Zwf[w], rcond=-1)[0]
    bc = np.linalg.lstsq(basis.T, wf[w], rcond=-1)[0]
    # as the reference, the fit uses batoid's values under Zemax's mask,
    # vignetted or not (a masked array's data)
    np.testing.assert_allclose(zc[4:], bc[4:], rtol=0, atol=0.01)


def _gram(x, y, w, jmax):
    """[Z_1 .. Z_jmax, w] Gram matrix in numpy (the device pass's output)."""
    C, nmax = zernike_table(jmax)
    cols = []
    for j in range(jmax):
        v = np.zeros_like(x)
        for d in range(nmax + 1):
            for q in range(d + 1):
                c = C[j, d * (d + 1) // 2 + q]
                if c:
                    v += c * x ** (d - q) * y ** q
        cols.append(v)
    B = np.stack(cols + [w], axis=1)
    return B.T @ B, B[:, :jmax]


def _reference_rank(B, w):
    """numpy lstsq's rank with the reference's rcond (analysis.py:566)."""
    return np.linalg.lstsq(B, w, rcond=None)[2]


def test_normal_equations_rank_rule():
    """The host solve of zernike(): rank-deficient fits raise RuntimeError as
    the reference's lstsq check does (analysis.py:566-570), full-rank fits
    equal lstsq's solution."""
    from batoid_amd.analysis import solve_normal_equations
    rng = np.random.default_rng(3)
    jmax = 22
    # fewer rows than coefficients
    for n in (1, 5, 21):
        r = np.sqrt(rng.uniform(0, 1, n))
        th = rng.uniform(0, 2 * np.pi, n)
        x, y = r * np.cos(th), r * np.sin(th)
        G, B = _gram(x, y, rng.normal(size=n), jmax)
        assert _reference_rank(B, x) < jmax
        with pytest.raises(RuntimeError):
            solve_normal_equations(G, jmax)
    # many rows on a line or on a circle: degenerate bases
    t = rng.uniform(-1, 1, 500)
    th = rng.uniform(0, 2 * np.pi, 500)
    for x, y in ((t, np.zeros_like(t)), (0.7 * np.cos(th), 0.7 * np.sin(th))):
        G, B = _gram(x, y, rng.normal(size=len(x)), jmax)
        assert _reference_rank(B, x) < jmax
        with pytest.raises(RuntimeError):
            solve_normal_equations(G, jmax)
    # a full disk: full rank, the lstsq solution
    r = np.sqrt(rng.uniform(0, 1, 2000))
    th = rng.uniform(0, 2 * np.pi, 2000)
    x, y = r * np.cos(th), r * np.sin(th)
    truth = rng.normal(size=jmax)
    _, B = _gram(x, y, np.zeros_like(x), jmax)
    w = B @ truth + 1e-3 * rng.normal(size=len(x))
    G, B = _gram(x, y, w, jmax)
    assert _reference_rank(B, w) == jmax
    got = solve_normal_equations(G, jmax)
    assert got[0] == 0.0
    np.testing.assert_allclose(got[1:], np.linalg.lstsq(B, w, rcond=None)[0],
                               rtol=0, atol=1e-9)
