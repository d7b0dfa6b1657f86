"""Benchmark optics and ray bundles.

* ``newtonian()``: one Paraboloid mirror (R = 10 m in batoid's frame, where
  rays arrive travelling -z; 2 m aperture, 0.4 m central obscuration) and a
  flat Detector at its focus z = R/2, in vacuum.
* ``tma()``: a 3-mirror anastigmat with three fused-silica lenses (Sellmeier),
  a flat filter (a Plane/Plane silica plate) and a detector -- the LSST
  r-band prescription (batoid/data/LSST/LSST_r.yaml:17-188) with the
  primary's even asphere extended to five terms and the filter's two
  shallow spheres (R = 5.632 / 5.606 m there) made flat.  12 interfaces.
* ``tma_rays()``: the hexapolar pupil of the TMA: asPolar(outer=4.18,
  inner=2.558, nrad=1120, naz=11080) = 9,998,544 rays.
* ``tma_fields()``: 189 field points uniform in a 1.75 degree disk.
"""
import numpy as np

from .parse import parse_optic
from .rayVector import RayVector

WAVELENGTHS = (350e-9, 490e-9, 630e-9, 770e-9, 910e-9, 1050e-9)

_SILICA = {'type': 'SellmeierMedium', 'B1': 0.696183020, 'B2': 0.407925877,
           'B3': 0.897464057, 'C1': 0.00467926519, 'C2': 0.0135122244,
           'C3': 97.9323636}


def _lens(name, z, front, back, r_clear, dz):
    return {'type': 'Lens', 'name': name, 'medium': _SILICA, 'coordSys': {'z': z},
            'items': [
                {'type': 'RefractiveInterface', 'name': name + '_entrance', 'surface': front,
                 'obscuration': {'type': 'ClearCircle', 'radius': r_clear}},
                {'type': 'RefractiveInterface', 'name': name + '_exit', 'surface': back,
                 'obscuration': {'type': 'ClearCircle', 'radius': r_clear},
                 'coordSys': {'z': dz}}]}


def tma_config(m1_coefs=(0.0, -1.381e-9, -4.3e-20, 2.1e-25, -1.5e-30)):
    return {
        'type': 'CompoundOptic', 'name': 'TMA', 'inMedium': 1.0,
        'backDist': 15.0, 'sphereRadius': 5.0, 'pupilSize': 8.36,
        'pupilObscuration': 0.612,
        'stopSurface': {'type': 'Interface', 'name': 'entrancePupil',
                        'surface': {'type': 'Plane'}, 'coordSys': {'z': 0.4393899}},
        'items': [
            {'type': 'Mirror', 'name': 'M1',
             'surface': {'type': 'Asphere', 'R': 19.835, 'conic': -1.215,
                         'coefs': list(m1_coefs)},
             'obscuration': {'type': 'ClearAnnulus', 'outer': 4.18, 'inner': 2.558}},
            {'type': 'Mirror', 'name': 'M2',
             'surface': {'type': 'Asphere', 'R': 6.788, 'conic': -0.222,
                         'coefs': [0.0, 1.274e-5, 9.68e-7]},
             'obscuration': {'type': 'ClearAnnulus', 'outer': 1.71, 'inner': 0.9},
             'coordSys': {'z': 6.1562006}},
            {'type': 'Mirror', 'name': 'M3',
             'surface': {'type': 'Asphere', 'R': 8.3445, 'conic': 0.155,
                         'coefs': [0.0, 4.5e-7, 8.15e-9]},
             'obscuration': {'type': 'ClearAnnulus', 'outer': 2.508, 'inner': 0.55},
             'coordSys': {'z': -0.2338}},
            {'type': 'CompoundOptic', 'name': 'Camera',
             'coordSys': {'z': 3.3974725882045593},
             'items': [
                 _lens('L1', 0.0, {'type': 'Sphere', 'R': 2.824},
                       {'type': 'Sphere', 'R': 5.021}, 0.775, 0.08223),
                 _lens('L2', 0.49487202, {'type': 'Plane'},
                       {'type': 'Asphere', 'R': 2.529, 'conic': -1.57,
                        'coefs': [0.0, -0.001656]}, 0.551, 0.03),
                 _lens('Filter', 0.87445202, {'type': 'Plane'}, {'type': 'Plane'},
                       0.375, 0.0179),
                 _lens('L3', 0.94345202, {'type': 'Quadric', 'R': 3.169, 'conic': -0.962},
                       {'type': 'Sphere', 'R': -13.36}, 0.361, 0.06),
                 {'type': 'Detector', 'name': 'Detector', 'surface': {'type': 'Plane'},
                  'obscuration': {'type': 'ClearCircle', 'radius': 0.4},
                  'coordSys': {'z': 1.03195202}},
             ]},
        ]}


def tma(**kwargs):
    return parse_optic(tma_config(**kwargs))


def tma_fields(n=189, seed=5772156, radius_deg=1.75):
    """(theta_x, theta_y) in radians, uniform in a disk."""
    rng = np.random.default_rng(seed)
    r = np.deg2rad(radius_deg) * np.sqrt(rng.uniform(size=n))
    th = rng.uniform(0, 2 * np.pi, size=n)
    return r * np.cos(th), r * np.sin(th)


def tma_rays(optic, theta_x, theta_y, wavelength, nrad=1120, naz=11080):
    return RayVector.asPolar(optic=optic, outer=4.18, inner=2.558, nrad=nrad, naz=naz,
                             theta_x=theta_x, theta_y=theta_y, wavelength=wavelength)


def newtonian_config():
    return {
        'type': 'CompoundOptic', 'name': 'Newtonian', 'inMedium': 1.0,
        'backDist': 40.0, 'pupilSize': 2.0, 'pupilObscuration': 0.2,
        'items': [
            {'type': 'Mirror', 'name': 'Primary', 'surface': {'type': 'Paraboloid', 'R': 10.0},
             'obscuration': {'type': 'ClearAnnulus', 'inner': 0.2, 'outer': 1.0}},
            {'type': 'Detector', 'name': 'Detector', 'surface': {'type': 'Plane'},
             'coordSys': {'z': 5.0}},
        ]}


def newtonian():
    return parse_optic(newtonian_config())


def newtonian_rays(optic, n=1_000_000, seed=57721, wavelength=500e-9, theta_deg=0.1):
    return RayVector.asPolar(optic=optic, outer=1.0, inner=0.0, nrandom=n,
                             rng=np.random.default_rng(seed), wavelength=wavelength,
                             theta_x=np.deg2rad(theta_deg), theta_y=0.0)


def perturbed_tma(jmax=28, zernike_rms=50e-9, zernike_seed=577215, grid=204,
                  grid_rms=10e-9, grid_seed=5772156649, smoothing=4.0, **kwargs):
    """The TMA with every mirror a Sum of its Asphere, a Zernike series
    (j <= 28, coefficients N(0, 50 nm), R_outer = the mirror's clear outer
    radius) and a Bicubic 204 x 204 grid over +-1.02 R_outer of Gaussian-
    smoothed white noise scaled to 10 nm rms."""
    from scipy.ndimage import gaussian_filter
    from .surface import Sum, Zernike, Bicubic
    optic = tma(**kwargs)
    zrng = np.random.default_rng(zernike_seed)
    brng = np.random.default_rng(grid_seed)
    for name in ("M1", "M2", "M3"):
        iface = optic[name]
        R = iface.R_outer
        coef = zrng.normal(0.0, zernike_rms, size=jmax + 1)
        coef[0] = 0.0
        xs = np.linspace(-1.02 * R, 1.02 * R, grid)
        noise = gaussian_filter(brng.normal(size=(grid, grid)), smoothing)
        noise *= grid_rms / np.std(noise)
        surf = Sum(iface.surface, Zernike(coef, R_outer=R), Bicubic(xs, xs, noise))
        optic = optic.withSurface(name, surf)
    return optic


def _asphere_sag(asph, x, y):
    """Host sag of an Asphere (quadric + sum_i alpha_i r^(2i)), for fitting."""
    r2 = x * x + y * y
    R, k = asph.R, asph.conic
    out = r2 / (R * (1 + np.sqrt(1 - (1 + k) * r2 / R ** 2)))
    for i, a in enumerate(asph._padded_coefs):
        out = out + a * r2 ** (i + 1)
    return out


def _fit_zernike(fun, R_outer, R_inner, jmax, npts=20000, seed=1):
    """Least-squares Zernike coefficients (Noll 1..jmax, circular basis over
    R_outer) of fun(x, y) sampled uniformly over the annulus."""
    from .zernike_poly import zernike_coef_arrays
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(R_inner ** 2, R_outer ** 2, npts))
    th = rng.uniform(0, 2 * np.pi, npts)
    x, y = r * np.cos(th), r * np.sin(th)
    cols = []
    for j in range(1, jmax + 1):
        e = np.zeros(jmax + 1)
        e[j] = 1.0
        c, _, _ = zernike_coef_arrays(e, R_outer, 0.0)
        cols.append(np.polynomial.polynomial.polyval2d(x, y, c))
    A = np.stack(cols, axis=1)
    sol, *_ = np.linalg.lstsq(A, fun(x, y), rcond=None)
    return np.concatenate([[0.0], sol])


def freeform_tma(jfit=45, **kwargs):
    """The perturbed TMA with freeform mirrors: each mirror is
    Sum(Zernike(fit of its asphere + the 50 nm perturbation, j <= 45),
    Bicubic) -- the same optical surfaces to ~1e-8 m, but with no conic term,
    so the first child's Newton starts from the ray origin (dt = 0) as the
    reference does for polynomial surfaces (surface.cpp:22-49)."""
    from .surface import Sum, Zernike
    base = perturbed_tma(**kwargs)
    optic = base
    for name in ("M1", "M2", "M3"):
        iface = base[name]
        asph, zern, bic = iface.surface.surfaces
        R = iface.R_outer
        fit = _fit_zernike(lambda x, y: _asphere_sag(asph, x, y), R, iface.R_inner, jfit)
        coef = fit.copy()
        coef[:len(zern.coef)] += zern.coef
        optic = optic.withSurface(name, Sum(Zernike(coef, R_outer=R), bic))
    return optic
