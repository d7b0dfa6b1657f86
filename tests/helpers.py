"""Shared factories for the parity tests: seeded rays, one instance of every
surface / medium / obscuration kind, and bit-exact comparisons."""
import numpy as np

import batoid_amd as bt

FCOLS = ("x", "y", "z", "vx", "vy", "vz", "t", "wavelength", "flux")
BCOLS = ("vignetted", "failed")

SILICA = (0.6961663, 0.4079426, 0.8974794, 0.00467914825849,
          0.013512063073959999, 97.93400253792099)          # tests/test_Medium.py:69-74
KBK7 = (2.2705778, -0.010059376, 0.010414999, 0.00028872517, -2.2214495e-5,
        1.4258559e-6)                                        # tests/test_Medium.py:108-113


def bicubic_surface(seed=5, n=48, half=1.3, nanpolicy='nan'):
    """A smooth asphere-like grid surface with small seeded ripples."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-half, half, n)
    ys = np.linspace(-half, half, n + 3)
    xx, yy = np.meshgrid(xs, ys)
    zz = (xx**2 + yy**2) / (2 * 7.0) + 1e-4 * np.sin(3 * xx + 1) * np.cos(2 * yy)
    zz += 1e-7 * rng.normal(size=zz.shape)
    return bt.Bicubic(xs, ys, zz, nanpolicy=nanpolicy)


def zernike_surface(seed=57, jmax=22, R_outer=1.2, R_inner=0.0):
    rng = np.random.default_rng(seed)
    coef = rng.normal(scale=1e-4, size=jmax + 1)
    return bt.Zernike(coef, R_outer=R_outer, R_inner=R_inner)


def surfaces():
    """(name, surface) for every kind, including a three-child Sum."""
    return [
        ("plane", bt.Plane()),
        ("tilted", bt.Tilted(0.01, -0.02)),
        ("sphere+", bt.Sphere(2.3)),
        ("sphere-", bt.Sphere(-3.1)),
        ("paraboloid+", bt.Paraboloid(4.5)),
        ("paraboloid-", bt.Paraboloid(-4.5)),
        ("quadric", bt.Quadric(6.1, -0.7)),
        ("quadric2", bt.Quadric(-8.0, 0.4)),
        ("asphere_m1", bt.Asphere(19.835, -1.215, [0.0, -1.381e-9])),
        ("asphere_l2", bt.Asphere(2.529, -1.57, [0.0, -0.001656])),
        ("asphere_r2", bt.Asphere(5.0, 0.2, [1e-3, -2e-4, 3e-6], imin=1)),
        ("bicubic", bicubic_surface()),
        ("bicubic_zero", bicubic_surface(nanpolicy='zero')),
        ("zernike", zernike_surface()),
        ("zernike_annular", zernike_surface(R_inner=0.4)),
        ("sum", bt.Sum(bt.Asphere(9.0, -0.9, [1e-4, -2e-6]), zernike_surface(seed=58),
                       bicubic_surface(seed=6))),
        ("sum_plane", bt.Sum(bt.Plane(), bt.Paraboloid(12.0))),
        # Sums of Sums (sum.cpp evaluates each child, Sum or not, by its own
        # sag / normal / timeToIntersect), nested to the three levels the
        # GPU evaluates, and a Sum of eleven children
        ("sum_nested", bt.Sum(bt.Sum(bt.Asphere(9.0, -0.9, [1e-4]), zernike_surface(seed=59)),
                              bt.Sum(bicubic_surface(seed=7), bt.Tilted(0.001, 0.0005)),
                              bt.Plane())),
        ("sum_deep", bt.Sum(bt.Sum(bt.Sum(bt.Sum(bt.Paraboloid(8.0), bicubic_surface(seed=8)),
                                          bt.Tilted(-0.002, 0.001)),
                                   zernike_surface(seed=60, jmax=11)),
                            bt.Sphere(40.0))),
        ("sum_wide", bt.Sum([bt.Quadric(7.0, -0.4), bicubic_surface(seed=9)]
                            + [zernike_surface(seed=61 + k, jmax=6) for k in range(8)]
                            + [bt.Tilted(0.0005, -0.0007)])),
    ]


def media():
    return [
        ("const", bt.ConstMedium(1.46)),
        ("table", bt.TableMedium([300e-9, 500e-9, 700e-9, 1100e-9], [1.5, 1.47, 1.46, 1.45])),
        ("sellmeier", bt.SellmeierMedium(*SILICA)),
        ("sumita", bt.SumitaMedium(*KBK7)),
        ("air", bt.Air()),
    ]


def obscurations():
    return [
        ("circle", bt.ObscCircle(0.6, 0.1, -0.05)),
        ("annulus", bt.ObscAnnulus(0.2, 0.7)),
        ("rectangle", bt.ObscRectangle(0.5, 0.3, 0.1, 0.05, theta=0.3)),
        ("ray", bt.ObscRay(0.1, 1.1, 0.02, -0.03)),
        ("polygon", bt.ObscPolygon([0.0, 0.5, 0.4, -0.3, -0.5], [-0.5, -0.2, 0.6, 0.5, 0.0])),
        ("clear_annulus", bt.ObscNegation(bt.ObscAnnulus(0.25, 0.9))),
        ("union", bt.ObscUnion(bt.ObscCircle(0.2), bt.ObscRay(0.05, 0.0),
                               bt.ObscRay(0.05, np.pi / 2))),
        ("intersection", bt.ObscIntersection(bt.ObscNegation(bt.ObscCircle(0.3)),
                                             bt.ObscCircle(0.8))),
        ("nested", bt.ObscNegation(bt.ObscUnion(
            bt.ObscIntersection(bt.ObscCircle(0.9), bt.ObscNegation(bt.ObscCircle(0.2))),
            bt.ObscRectangle(0.1, 2.0)))),
    ]


def random_rays(n, seed=11, z0=1.0, n_index=1.0, spread=0.15, radius=1.0,
                wl=(400e-9, 1000e-9), prefail=0.01, previg=0.02, upward=0.0):
    """Rays starting at height z0 heading roughly along -z (a fraction
    ``upward`` along +z), with some pre-failed and pre-vignetted flags."""
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    x = r * np.cos(th)
    y = r * np.sin(th)
    z = np.full(n, z0) + rng.uniform(-0.01, 0.01, n)
    vx = rng.normal(scale=spread, size=n)
    vy = rng.normal(scale=spread, size=n)
    vz = -np.ones(n)
    up = rng.uniform(size=n) < upward
    vz[up] = 1.0
    norm = np.sqrt(vx * vx + vy * vy + vz * vz) * n_index
    return dict(
        x=x, y=y, z=z, vx=vx / norm, vy=vy / norm, vz=vz / norm,
        t=rng.uniform(0, 1, n), wavelength=rng.uniform(wl[0], wl[1], n),
        flux=rng.uniform(0.5, 1.5, n),
        vignetted=rng.uniform(size=n) < previg, failed=rng.uniform(size=n) < prefail)


def transforms():
    """(dr, drot) pairs: identity and a tilted, shifted frame."""
    ident = (np.zeros(3), np.eye(3))
    cs = bt.CoordSys().shiftGlobal([0.01, -0.02, 0.05]).rotateLocal(
        bt.RotX(0.01) @ bt.RotY(-0.02))
    ct = bt.CoordTransform(bt.CoordSys(), cs)
    return [("identity", ident), ("tilted", (ct.dr, ct.drot))]


def assert_cols_identical(a, b, cols=FCOLS + BCOLS, what=""):
    """Bit-for-bit equality, NaN in the same places counting as equal."""
    for k in cols:
        x = np.asarray(a[k])
        y = np.asarray(b[k])
        if k in BCOLS:
            np.testing.assert_array_equal(x.astype(bool), y.astype(bool), err_msg=f"{what} {k}")
        else:
            same = (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))
            if not same.all():
                i = np.flatnonzero(~same)[:5]
                raise AssertionError(f"{what} column {k}: {(~same).sum()} of {len(x)} differ; "
                                     f"first {i}: {x[i]} vs {y[i]}")
