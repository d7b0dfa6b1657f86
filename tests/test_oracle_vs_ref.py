"""Pin the C restatement (oracle/oracle.c) bit-for-bit against the
reference's own C++ core compiled from its sources (oracle/_ref).

Same seeded inputs through both, for every surface kind x interaction x
frame, every medium, every obscuration kind and the RayVector field
operations.  Skipped where oracle/_ref has not been built."""
import ctypes

import numpy as np
import pytest

import oracle
from oracle.host import HostRays, host_records, nthreads
from helpers import (surfaces, media, obscurations, random_rays, transforms,
                     assert_cols_identical, FCOLS)

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")

N = 4000
NITER = 5


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
@pytest.mark.parametrize("action", ["intersect", "reflect", "refract"])
def test_surface_interactions(name, surf, action):
    O, R = oracle.load(), oracle.load_ref()
    tab = host_records(surf)
    med = host_records(media()[2][1])
    for tname, (dr, drot) in transforms():
        for seed, upward in ((1, 0.0), (2, 0.3)):
            base = HostRays(**random_rays(N, seed=seed, upward=upward, spread=0.3))
            a, b = base.copy(), base.copy()
            args = (tab.ptr(), _dv(dr), _dv(drot))
            if action == "refract":
                getattr(O, "or_refract")(*args, med.ptr(), a.struct, NITER, nthreads())
                getattr(R, "ref_refract")(*args, med.ptr(), b.struct, NITER, 1)
            else:
                getattr(O, f"or_{action}")(*args, a.struct, NITER, nthreads())
                getattr(R, f"ref_{action}")(*args, b.struct, NITER, 1)
            assert_cols_identical(a, b, what=f"{name}/{action}/{tname}/{seed}")
            # the case must exercise both success and failure
            if name not in ("plane", "tilted", "sum_plane"):
                assert 0 < a["failed"].sum() < N


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
def test_coated_interactions_and_rsplit(name, surf):
    """Coated intersect / reflect / refract and rSplit (batoid.cpp:175-674
    with a SimpleCoating) against the reference, every surface, two frames,
    failures, total internal reflection and pre-failed rays included."""
    O, R = oracle.load(), oracle.load_ref()
    tab = host_records(surf)
    med = host_records(media()[2][1])
    refl, trans = 0.0371, 0.9588
    for tname, (dr, drot) in transforms()[:2]:
        base = HostRays(**random_rays(N, seed=9, upward=0.3, spread=0.3))
        args = (tab.ptr(), _dv(dr), _dv(drot))
        for action, coef in ((0, trans), (1, refl), (2, trans)):
            a, b = base.copy(), base.copy()
            if action == 2:
                O.or_refract_coated(*args, med.ptr(), coef, a.struct, NITER, nthreads())
            else:
                getattr(O, "or_intersect_coated" if action == 0 else "or_reflect_coated")(
                    *args, coef, a.struct, NITER, nthreads())
            R.ref_coated(action, *args, med.ptr(), refl, trans, b.struct, NITER, 1)
            assert_cols_identical(a, b, what=f"{name}/coated{action}/{tname}")
        a, b = base.copy(), base.copy()
        sa, sb = base.copy(), base.copy()
        O.or_rsplit(*args, med.ptr(), refl, trans, a.struct, sa.struct, NITER, nthreads())
        R.ref_rsplit(*args, med.ptr(), refl, trans, b.struct, sb.struct, NITER, 1)
        assert_cols_identical(a, b, what=f"{name}/rsplit refracted/{tname}")
        assert_cols_identical(sa, sb, what=f"{name}/rsplit reflected/{tname}")
        ok = (base["failed"] == 0) & (a["failed"] == 0)
        assert ok.sum() > 0
        np.testing.assert_array_equal(a["flux"][ok], base["flux"][ok] * trans)


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
def test_sag_normal(name, surf):
    O, R = oracle.load(), oracle.load_ref()
    tab = host_records(surf)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.5, 1.5, 5000)
    y = rng.uniform(-1.5, 1.5, 5000)
    for fn, width in (("sag_n", 1), ("normal_n", 3)):
        oa = np.empty(width * len(x))
        ra = np.empty(width * len(x))
        getattr(O, "or_" + fn)(tab.ptr(), x.ctypes.data, y.ctypes.data, oa.ctypes.data,
                               len(x), *(() if False else (nthreads(),)))
        getattr(R, "ref_" + fn)(tab.ptr(), x.ctypes.data, y.ctypes.data, ra.ctypes.data, len(x))
        assert_cols_identical({"v": oa}, {"v": ra}, cols=("v",), what=f"{name} {fn}")


@pytest.mark.parametrize("name,med", media(), ids=[m[0] for m in media()])
def test_media(name, med):
    O, R = oracle.load(), oracle.load_ref()
    tab = host_records(med)
    w = np.random.default_rng(4).uniform(250e-9, 1200e-9, 3000)
    oa, ra = np.empty_like(w), np.empty_like(w)
    O.or_get_n_n(tab.ptr(), w.ctypes.data, oa.ctypes.data, len(w), nthreads())
    R.ref_get_n_n(tab.ptr(), w.ctypes.data, ra.ctypes.data, len(w))
    assert_cols_identical({"n": oa}, {"n": ra}, cols=("n",), what=name)


@pytest.mark.parametrize("name,obsc", obscurations(), ids=[o[0] for o in obscurations()])
def test_obscurations(name, obsc):
    O, R = oracle.load(), oracle.load_ref()
    tab = host_records(obsc)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 20000)
    y = rng.uniform(-1, 1, 20000)
    oa = np.empty(len(x), np.uint8)
    ra = np.empty(len(x), np.uint8)
    O.or_contains_n(tab.ptr(), tab.count, x.ctypes.data, y.ctypes.data, oa.ctypes.data,
                    len(x), nthreads())
    R.ref_contains_n(tab.ptr(), tab.count, x.ctypes.data, y.ctypes.data, ra.ctypes.data, len(x))
    np.testing.assert_array_equal(oa, ra)
    assert 0 < oa.sum() < len(x)
    base = HostRays(**random_rays(5000, seed=6))
    a, b = base.copy(), base.copy()
    O.or_obscure(tab.ptr(), tab.count, a.struct, nthreads())
    R.ref_obscure(tab.ptr(), tab.count, b.struct, 1)
    assert_cols_identical(a, b, what=name)


def test_transforms_and_finish():
    O, R = oracle.load(), oracle.load_ref()
    for tname, (dr, drot) in transforms():
        base = HostRays(**random_rays(3000, seed=7))
        for fwd in ("forward", "reverse"):
            a, b = base.copy(), base.copy()
            getattr(O, f"or_apply_{fwd}_transform")(_dv(dr), _dv(drot), a.struct, nthreads())
            getattr(R, f"ref_apply_{fwd}_transform")(_dv(dr), _dv(drot), b.struct, 1)
            assert_cols_identical(a, b, what=f"{tname} {fwd}")
        xa, ya, za = (base[k].copy() for k in "xyz")
        xb, yb, zb = (base[k].copy() for k in "xyz")
        vv = np.array([0.01, -0.02, -0.99])
        O.or_finish_parallel(_dv(dr), _dv(drot), _dv(vv), xa.ctypes.data, ya.ctypes.data,
                             za.ctypes.data, len(xa))
        R.ref_finish_parallel(_dv(dr), _dv(drot), _dv(vv), xb.ctypes.data, yb.ctypes.data,
                              zb.ctypes.data, len(xb))
        assert_cols_identical({"x": xa, "y": ya, "z": za}, {"x": xb, "y": yb, "z": zb},
                              cols=("x", "y", "z"))


def test_phase_and_sum_amplitude():
    O, R = oracle.load(), oracle.load_ref()
    base = HostRays(**random_rays(4000, seed=8))
    oa = np.empty(len(base))
    ra = np.empty(len(base))
    O.or_phase(base.struct, 0.1, -0.2, 0.3, 1.5, oa.ctypes.data, nthreads())
    R.ref_phase(base.struct, 0.1, -0.2, 0.3, 1.5, ra.ctypes.data, 1)
    assert_cols_identical({"p": oa}, {"p": ra}, cols=("p",))
    for ign in (0, 1):
        so = np.empty(2)
        sr = np.empty(2)
        O.or_sum_amplitude(base.struct, 0.1, -0.2, 0.3, 1.5, ign, so.ctypes.data, 1)
        R.ref_sum_amplitude(base.struct, 0.1, -0.2, 0.3, 1.5, ign, sr.ctypes.data, 1)
        np.testing.assert_array_equal(so, sr)
