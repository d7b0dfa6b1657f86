"""GPU ghost splitting: coated interactions, rSplit, the ray selection and
traceSplit, against the oracle (bit-for-bit) and the reference's own
traceSplit tests (tests/test_rsplit.py)."""
import ctypes
import os

import numpy as np
import pytest
import torch

import batoid_amd as bt
import oracle
from batoid_amd import _lib
from oracle import split as OS
from oracle.host import HostRays, host_records, nthreads
from helpers import (surfaces, media, random_rays, transforms, assert_cols_identical,
                     FCOLS, BCOLS)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
N = 20000


def _gpu(h, coordSys=bt.globalCoordSys):
    rv = bt.RayVector(**{k: h[k] for k in FCOLS}, vignetted=h["vignetted"].astype(bool),
                      failed=h["failed"].astype(bool))
    rv.coordSys = coordSys
    return rv


def _host(rv):
    torch.cuda.synchronize()
    return {k: getattr(rv, k) for k in FCOLS + BCOLS}


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


def _frame(dr, drot):
    return bt.CoordSys(origin=dr, rot=np.asarray(drot).reshape(3, 3))


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
def test_coated_and_rsplit_bit_exact(name, surf):
    O = oracle.load()
    tab = host_records(surf)
    med = media()[2][1]
    mtab = host_records(med)
    coat = bt.SimpleCoating(0.0371, 0.9588)
    for tname, (dr, drot) in transforms():
        cs = _frame(dr, drot)
        base = HostRays(**random_rays(N, seed=21, upward=0.3, spread=0.3))
        args = (tab.ptr(), _dv(dr), _dv(drot))
        # coated interactions
        for action in ("intersect", "reflect", "refract"):
            ref = base.copy()
            if action == "refract":
                O.or_refract_coated(*args, mtab.ptr(), coat.transmissivity, ref.struct, 5,
                                    nthreads())
                rv = surf.refract(_gpu(base), bt.ConstMedium(1.0), med, coordSys=cs,
                                  coating=coat)
            elif action == "reflect":
                O.or_reflect_coated(*args, coat.reflectivity, ref.struct, 5, nthreads())
                rv = surf.reflect(_gpu(base), coordSys=cs, coating=coat)
            else:
                O.or_intersect_coated(*args, coat.transmissivity, ref.struct, 5, nthreads())
                rv = surf.intersect(_gpu(base), coordSys=cs, coating=coat)
            assert_cols_identical(_host(rv), ref.cols, what=f"{name}/{action}/{tname}")
        # rSplit
        a, sa = base.copy(), base.copy()
        O.or_rsplit(*args, mtab.ptr(), coat.reflectivity, coat.transmissivity, a.struct,
                    sa.struct, 5, nthreads())
        refr, refl = surf.rSplit(_gpu(base), bt.ConstMedium(1.0), med, coat, coordSys=cs)
        assert_cols_identical(_host(refr), a.cols, what=f"{name}/rsplit refracted/{tname}")
        assert_cols_identical(_host(refl), sa.cols, what=f"{name}/rsplit reflected/{tname}")
        assert refr.coordSys == cs and refl.coordSys == cs


@pytest.mark.parametrize("name,surf", surfaces()[:4], ids=[s[0] for s in surfaces()[:4]])
def test_rsplit_with_obscuration_bit_exact(name, surf):
    """bt_rsplit with the obscuration folded in equals rSplit followed by
    the obscuration of both parts (optic.py:491-494), bit for bit."""
    O = oracle.load()
    tab = host_records(surf)
    med = media()[2][1]
    mtab = host_records(med)
    coat = bt.SimpleCoating(0.0371, 0.9588)
    obsc = bt.ObscNegation(bt.ObscAnnulus(0.2, 0.8))
    otab = host_records(obsc)
    for tname, (dr, drot) in transforms():
        cs = _frame(dr, drot)
        base = HostRays(**random_rays(N, seed=23, upward=0.3, spread=0.3))
        a, sa = base.copy(), base.copy()
        O.or_rsplit(tab.ptr(), _dv(dr), _dv(drot), mtab.ptr(), coat.reflectivity,
                    coat.transmissivity, a.struct, sa.struct, 5, nthreads())
        O.or_obscure(otab.ptr(), otab.count, a.struct, nthreads())
        O.or_obscure(otab.ptr(), otab.count, sa.struct, nthreads())
        refr, refl = surf.rSplit(_gpu(base), bt.ConstMedium(1.0), med, coat, coordSys=cs,
                                 obscuration=obsc)
        assert_cols_identical(_host(refr), a.cols, what=f"{name}/rsplit+obsc refracted/{tname}")
        assert_cols_identical(_host(refl), sa.cols, what=f"{name}/rsplit+obsc reflected/{tname}")


@pytest.mark.parametrize("name,surf", [surfaces()[i] for i in (2, 9, 15)],
                         ids=[surfaces()[i][0] for i in (2, 9, 15)])
def test_rsplit_live_skips_dropped_rays(name, surf):
    """bt_rsplit_live on a set that still holds rays traceSplit's filter
    dropped (vignetted, flux below minFlux, NaN flux): those rays are left
    exactly as they were and their split part is marked vignetted; every
    kept ray is split bit for bit as bt_rsplit splits it (the analytic-kinds
    kernel for the sphere and the asphere, the all-kinds one for the Sum),
    except that a reflected ray the filter drops is written only as its
    flux and flags; and the kernel's counts of kept output rays are the
    filter's."""
    med = media()[2][1]
    coat = bt.SimpleCoating(0.0371, 0.9588)
    obsc = bt.ObscNegation(bt.ObscAnnulus(0.2, 0.8))
    rng = np.random.default_rng(29)
    minFlux = 1e-3
    for tname, (dr, drot) in transforms():
        cs = _frame(dr, drot)
        base = HostRays(**random_rays(N, seed=27, upward=0.3, spread=0.3))
        # reflected rays (x 0.0371) pass minFlux from flux 0.027 up
        flux = rng.uniform(0.0, 3e-2, N)
        flux[rng.uniform(size=N) < 0.02] = np.nan
        base.cols["flux"] = flux
        base.cols["vignetted"] = (rng.uniform(size=N) < 0.2).astype(np.uint8)
        keep = (base["vignetted"] == 0) & (base["flux"] >= minFlux)
        assert 0 < keep.sum() < N
        full_r, full_s = surf.rSplit(_gpu(base), bt.ConstMedium(1.0), med, coat, coordSys=cs,
                                     obscuration=obsc)
        counts = torch.zeros(2 * 64, dtype=torch.int64, device="cuda")
        live_r, live_s = surf.rSplit(_gpu(base), bt.ConstMedium(1.0), med, coat, coordSys=cs,
                                     obscuration=obsc, _live_min_flux=minFlux,
                                     _live_counts=counts)
        fr, fs, lr, ls = (_host(r) for r in (full_r, full_s, live_r, live_s))
        # the kernel's own counts of the rays each output keeps
        for c, h in zip(counts.view(2, 64).sum(1).tolist(), (fr, fs)):
            assert c == int(((h["vignetted"] == 0) & (h["flux"] >= minFlux) & keep).sum())
        skept = keep & (fs["vignetted"] == 0) & (fs["flux"] >= minFlux)
        sdrop = keep & ~skept
        assert skept.sum() > 0 and sdrop.sum() > 0
        for k in FCOLS + BCOLS:
            assert_cols_identical({k: lr[k][keep]}, {k: fr[k][keep]}, cols=(k,),
                                  what=f"{name}/{tname} kept refracted")
            assert_cols_identical({k: ls[k][skept]}, {k: fs[k][skept]}, cols=(k,),
                                  what=f"{name}/{tname} kept reflected")
            assert_cols_identical({k: lr[k][~keep]}, {k: base[k][~keep]}, cols=(k,),
                                  what=f"{name}/{tname} dropped rays untouched")
        for k in ("flux",) + BCOLS:
            assert_cols_identical({k: ls[k][sdrop]}, {k: fs[k][sdrop]}, cols=(k,),
                                  what=f"{name}/{tname} reflected rays the filter drops")
        assert ls["vignetted"][~keep].all()
        # the filter after the split keeps the same rays either way
        for a, b in ((live_r, full_r), (live_s, full_s)):
            sa, sb = _host(a._select(minFlux)), _host(b._select(minFlux))
            cnt = len(sa["x"])
            assert cnt <= len(sb["x"])
            kb = (fr if b is full_r else fs)
            kept_b = (kb["vignetted"] == 0) & (kb["flux"] >= minFlux) & keep
            assert cnt == int(kept_b.sum())
            assert_cols_identical(sa, {k: kb[k][kept_b] for k in FCOLS + BCOLS},
                                  what=f"{name}/{tname} filtered outputs")


def test_rsplit_live_analytic_flag_on_other_kinds_fails_rays():
    """bt_rsplit_live with kinds = BT_PLAN_ANALYTIC on a Sum: the
    analytic-kinds kernel cannot evaluate it, so every kept ray fails (and
    is vignetted) in both parts instead of being traced wrongly."""
    from batoid_amd import _device
    surf = dict(surfaces())["sum"]
    assert not surf._analytic
    med = media()[2][1]
    coat = bt.SimpleCoating(0.0371, 0.9588)
    base = HostRays(**random_rays(N, seed=35, upward=0.3, spread=0.3))
    rv = _gpu(base)
    split = bt.RayVector._empty(len(rv), rv.coordSys)
    ct = bt.CoordTransform(rv.coordSys, rv.coordSys)
    _lib.call("bt_rsplit_live", surf._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(ct.drot.ravel()), med._dev.ptr(), ctypes.byref(coat._record), None, 0,
              rv._rays, split._rays, 0.0, 5, _lib.PLAN_ANALYTIC, None, _device.stream())
    r, sp = _host(rv), _host(split)
    live = (base["vignetted"] == 0) & (base["failed"] == 0)
    assert live.any()
    assert r["failed"][live].all() and r["vignetted"][live].all()
    assert sp["failed"][live].all() and sp["vignetted"][live].all()


@pytest.mark.parametrize("action", ["intersect", "reflect", "refract"])
@pytest.mark.parametrize("name,surf", [surfaces()[i] for i in (2, 9, 15)],
                         ids=[surfaces()[i][0] for i in (2, 9, 15)])
def test_interact_live_skips_dropped_rays(name, surf, action):
    """bt_interact_live (traceSplit's Mirror and Baffle steps on carried
    sets): dropped rays untouched, kept rays bit-identical to the plain
    call, and the kernel's count of kept output rays the filter's."""
    from batoid_amd import trace as T
    coat = bt.SimpleCoating(0.3, 0.7)
    obsc = bt.ObscNegation(bt.ObscAnnulus(0.2, 0.8))
    rng = np.random.default_rng(31)
    minFlux = 1e-3
    med = media()[2][1]
    if action == "refract":
        def fn(surf, rv, cs, coating, obscuration, _live_min_flux=None, _live_counts=None):
            if _live_min_flux is None:
                return T.refract(surf, rv, bt.vacuum, med, cs, coating=coating,
                                 obscuration=obscuration)
            ct = bt.CoordTransform(rv.coordSys, cs)
            T._interact_live(_lib.ACT_REFRACT, surf, rv, ct, med, coating, obscuration,
                             _live_min_flux, _live_counts)
            rv.coordSys = cs
            return rv
    else:
        fn = getattr(T, action)
    for tname, (dr, drot) in transforms():
        cs = _frame(dr, drot)
        base = HostRays(**random_rays(N, seed=33, upward=0.3, spread=0.3))
        # reflected rays (x 0.0371) pass minFlux from flux 0.027 up
        flux = rng.uniform(0.0, 3e-2, N)
        flux[rng.uniform(size=N) < 0.02] = np.nan
        base.cols["flux"] = flux
        base.cols["vignetted"] = (rng.uniform(size=N) < 0.2).astype(np.uint8)
        keep = (base["vignetted"] == 0) & (base["flux"] >= minFlux)
        full = _host(fn(surf, _gpu(base), cs, coating=coat, obscuration=obsc))
        counts = torch.zeros(64, dtype=torch.int64, device="cuda")
        live = _host(fn(surf, _gpu(base), cs, coating=coat, obscuration=obsc,
                        _live_min_flux=minFlux, _live_counts=counts))
        for k in FCOLS + BCOLS:
            assert_cols_identical({k: live[k][keep]}, {k: full[k][keep]}, cols=(k,),
                                  what=f"{name}/{action}/{tname} kept rays")
            assert_cols_identical({k: live[k][~keep]}, {k: base[k][~keep]}, cols=(k,),
                                  what=f"{name}/{action}/{tname} dropped rays untouched")
        kept_out = (full["vignetted"] == 0) & (full["flux"] >= minFlux) & keep
        assert int(counts.sum()) == int(kept_out.sum())


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 100_003, 2_000_000])
def test_select_matches_numpy(n):
    rng = np.random.default_rng(n)
    h = HostRays(**random_rays(max(n, 1), seed=n % 97))
    if n == 0:
        rv = bt.RayVector._empty(0, bt.globalCoordSys)
        assert len(rv._select(1e-3)) == 0
        return
    flux = rng.uniform(0, 2e-3, n)
    flux[rng.uniform(size=n) < 0.01] = np.nan
    h.cols["flux"] = flux
    h.cols["vignetted"] = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    rv = _gpu(h)
    for mf in (1e-3, 0.0, 5e-3):
        out = rv._select(mf)
        keep = (h["vignetted"] == 0) & (h["flux"] >= mf)
        assert len(out) == int(keep.sum())
        got = _host(out)
        for k in FCOLS + BCOLS:
            np.testing.assert_array_equal(got[k], h[k][keep].astype(got[k].dtype), err_msg=k)


def test_rsplit_reference_test():
    """tests/test_rsplit.py:8-34: rSplit equals reflect / refract with the
    same coating (up to the rounding of the reflection formula)."""
    rng = np.random.default_rng(5)
    for _ in range(30):
        R = rng.normal(0.7, 0.8)
        conic = rng.uniform(-2.0, 1.0)
        ncoef = int(rng.integers(0, 5))
        coefs = [rng.normal(0, 1e-10) for _ in range(ncoef)]
        asphere = bt.Asphere(R, conic, coefs)
        rays = bt.RayVector.asPolar(backDist=10.0, medium=bt.Air(), wavelength=500e-9,
                                    outer=0.25 * R, theta_x=rng.normal(0.0, 1e-8),
                                    theta_y=rng.normal(0.0, 1e-8), nrad=10, naz=10)
        coating = bt.SimpleCoating(0.9, 0.1)
        reflected = asphere.reflect(rays.copy(), coating=coating)
        m1, m2 = bt.Air(), bt.ConstMedium(1.1)
        refracted = asphere.refract(rays.copy(), m1, m2, coating=coating)
        refracted2, reflected2 = asphere.rSplit(rays, m1, m2, coating)
        for a, b in ((reflected, reflected2), (refracted, refracted2)):
            for k in ("x", "y", "z", "vx", "vy", "vz", "t", "flux"):
                np.testing.assert_allclose(getattr(a, k), getattr(b, k), rtol=0, atol=1e-14,
                                           equal_nan=True)
            np.testing.assert_array_equal(a.vignetted, b.vignetted)


def _simple_telescope():
    """tests/test_rsplit.py:38-80."""
    telescope = bt.CompoundOptic(
        name="simple", stopSurface=bt.Interface(bt.Plane()),
        items=[
            bt.Lens(name="L1", items=[
                bt.RefractiveInterface(bt.Plane(), name="L1_entrance",
                                       coordSys=bt.CoordSys(origin=[0, 0, 0.3]),
                                       inMedium=bt.ConstMedium(1.0),
                                       outMedium=bt.ConstMedium(1.1)),
                bt.RefractiveInterface(bt.Plane(), name="L1_exit",
                                       coordSys=bt.CoordSys(origin=[0, 0, 0.2]),
                                       inMedium=bt.ConstMedium(1.1),
                                       outMedium=bt.ConstMedium(1.0))]),
            bt.Mirror(bt.Plane(), name="Mirror"),
            bt.Detector(bt.Plane(), name="detector",
                        coordSys=bt.CoordSys(origin=[0, 0, 0.1]))],
        pupilSize=1.0, backDist=1.0, inMedium=bt.ConstMedium(1.0))
    for name in ("L1_entrance", "L1_exit"):
        telescope[name].forwardCoating = bt.SimpleCoating(0.02, 0.98)
        telescope[name].reverseCoating = bt.SimpleCoating(0.02, 0.98)
    telescope['detector'].forwardCoating = bt.SimpleCoating(0.02, 0.98)
    return telescope


def _check_paths(optic, rays, rF, rR, tol):
    """Every traceSplit output equals trace(path=) and the last entry of
    traceFull(path=) on its unvignetted rays (tests/test_rsplit.py:89-134)."""
    for r in list(rF) + list(rR):
        r2 = optic.trace(rays.copy(), path=r.path)
        w = ~r2.vignetted
        np.testing.assert_allclose(r.r, r2.r[w], atol=tol, rtol=0)
        np.testing.assert_allclose(r.v, r2.v[w], atol=tol, rtol=0)
        np.testing.assert_allclose(r.t, r2.t[w], atol=tol, rtol=0)
        tf = optic.traceFull(rays.copy(), path=r.path)
        keys = []
        for item in r.path:
            j = 0
            key = f"{item}_{j}"
            while key in keys:
                j += 1
                key = f"{item}_{j}"
            keys.append(key)
        assert keys == list(tf.keys())
        r3 = tf[keys[-1]]['out']
        w = ~r3.vignetted
        np.testing.assert_allclose(r.r, r3.r[w], atol=tol, rtol=0)
        np.testing.assert_allclose(r.v, r3.v[w], atol=tol, rtol=0)
        np.testing.assert_allclose(r.t, r3.t[w], atol=tol, rtol=0)


def _check_oracle(optic, rays, rF, rR, minFlux):
    """The same split through the oracle's restatement: the same paths in
    the same order, every column bit-identical."""
    oF, oR = OS.trace_split(optic, OS.Rays(HostRays.from_rv(rays), rays.coordSys), minFlux)
    for got, want in ((rF, oF), (rR, oR)):
        assert [g.path for g in got] == [w.path for w in want]
        for g, w in zip(got, want):
            assert g.coordSys == w.coordSys
            assert_cols_identical(_host(g), w.h.cols, what="/".join(g.path))


def test_traceSplit_simple():
    telescope = _simple_telescope()
    rays = bt.RayVector.asPolar(optic=telescope, wavelength=500e-9, theta_x=np.deg2rad(1.0),
                                theta_y=0.0, nrad=10, naz=60)
    rF, rR = telescope.traceSplit(rays.copy(), minFlux=1e-4)
    assert len(rF) > 1 and len(rR) > 1
    _check_paths(telescope, rays, rF, rR, 1e-14)
    _check_oracle(telescope, rays, rF, rR, 1e-4)


@pytest.mark.parametrize("name", ["DECam.yaml", "HSC.yaml", "LSST_r.yaml"])
def test_traceSplit_golden_optics(name):
    """tests/test_rsplit.py:137-213: refractive interfaces and the detector
    made 2 % reflective."""
    optic = bt.Optic.fromYaml(os.path.join(GOLD, "optics", name))
    for item in optic.itemDict.values():
        if isinstance(item, bt.RefractiveInterface):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
            item.reverseCoating = bt.SimpleCoating(0.02, 0.98)
        if isinstance(item, bt.Detector):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
    rays = bt.RayVector.asPolar(optic, wavelength=620e-9, theta_x=np.deg2rad(0.1),
                                theta_y=0.0, nrad=15, naz=90)
    minFlux = 1e-3
    rF, rR = optic.traceSplit(rays.copy(), minFlux=minFlux)
    assert len(rF) >= 1 and len(rF) + len(rR) > 2
    _check_paths(optic, rays, rF, rR, 1e-12)
    _check_oracle(optic, rays, rF, rR, minFlux)


def _coated_lsst():
    optic = bt.Optic.fromYaml(os.path.join(GOLD, "optics", "LSST_r.yaml"))
    for item in optic.itemDict.values():
        if isinstance(item, bt.RefractiveInterface):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
            item.reverseCoating = bt.SimpleCoating(0.02, 0.98)
        if isinstance(item, bt.Detector):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
    return optic


def test_traceSplit_deferred_equals_counted_rounds(monkeypatch):
    """traceSplit without count read-backs between rounds (sets dropped by
    their flux bound, the leaving sets counted once at the end) and with a
    read-back per round give the same sets in the same order, bit for
    bit."""
    optic = _coated_lsst()
    rays = bt.RayVector.asPolar(optic, wavelength=620e-9, theta_x=np.deg2rad(0.3),
                                theta_y=np.deg2rad(-0.2), nrad=30, naz=180)
    counted = []
    real = bt.RayVector._count_select

    def spy(rvs, minFlux):
        sel = real(rvs, minFlux)
        counted.append(sel.counts.tolist())
        return sel

    empties = 0
    for minFlux in (1e-3, 1e-4):
        assert optic._flux_bound(rays, minFlux) == 1.0
        counted.clear()
        with monkeypatch.context() as mp:
            mp.setattr(bt.RayVector, "_count_select", staticmethod(spy))
            dF, dR = optic.traceSplit(rays.copy(), minFlux=minFlux)
        # one count pass, at the end; the empty sets leave no output
        assert len(counted) == 1
        assert len(dF) + len(dR) == sum(c > 0 for c in counted[0])
        empties += counted[0].count(0)
        with monkeypatch.context() as mp:
            mp.setattr(bt.CompoundOptic, "_flux_bound", lambda self, rv, m: None)
            cF, cR = optic.traceSplit(rays.copy(), minFlux=minFlux)
        for got, want in ((dF, cF), (dR, cR)):
            assert [g.path for g in got] == [w.path for w in want]
            for g, w in zip(got, want):
                assert_cols_identical(_host(g), _host(w), what="/".join(g.path))
        assert len(dF) + len(dR) > 4
    # sets the flux bound kept but vignetting had emptied were traced on
    assert empties > 0


def test_traceSplit_flux_bound_preconditions():
    """The deferred traceSplit runs only where its recursion provably ends:
    minFlux > 0, every reflectivity that turns rays back below 1, a finite
    largest flux, and no failed ray the filter would keep."""
    optic = _coated_lsst()
    rays = bt.RayVector.asPolar(optic, wavelength=620e-9, theta_x=0.0, theta_y=0.0,
                                nrad=5, naz=30)
    i, j = (int(q) for q in np.flatnonzero(~np.asarray(rays.vignetted, bool))[[3, 5]])
    assert optic._flux_bound(rays, 1e-3) == 1.0
    assert optic._flux_bound(rays, 0.0) is None
    r = rays.copy()
    r.torch("flux")[i] = float("inf")
    assert optic._flux_bound(r, 1e-3) is None
    r = rays.copy()
    r.torch("failed")[i] = 1
    assert optic._flux_bound(r, 1e-3) is None
    r.torch("vignetted")[i] = 1                     # failed and vignetted: dropped anyway
    assert optic._flux_bound(r, 1e-3) == 1.0
    r = rays.copy()
    r.torch("flux")[i] = 7.0
    r.torch("flux")[j] = float("nan")
    assert optic._flux_bound(r, 1e-3) == 7.0
    item = next(i for i in optic.itemDict.values() if isinstance(i, bt.RefractiveInterface))
    item.reverseCoating = bt.SimpleCoating(1.0, 0.0)
    assert optic._flux_bound(rays, 1e-3) is None


def test_rsplit_live_without_split_part():
    """bt_rsplit_live asked for no split part (every split column NULL,
    n = 0): the refracted part is bit-identical to the call with one; a
    split with only some columns NULL is still refused."""
    from batoid_amd import trace as T
    from batoid_amd import _device
    surf = dict(surfaces())["asphere_l2"]
    med = media()[2][1]
    coat = bt.SimpleCoating(0.02, 0.98)
    rng = np.random.default_rng(41)
    base = HostRays(**random_rays(N, seed=43, upward=0.3, spread=0.3))
    base.cols["flux"] = rng.uniform(0.0, 3e-2, N)
    minFlux = 1e-3
    with_r, with_s = T.rSplit(surf, _gpu(base), bt.ConstMedium(1.0), med, coat,
                              _live_min_flux=minFlux)
    none_r, none_s = T.rSplit(surf, _gpu(base), bt.ConstMedium(1.0), med, coat,
                              _live_min_flux=minFlux, _no_split=True)
    assert none_s is None and with_s is not None
    assert_cols_identical(_host(none_r), _host(with_r), what="refracted part")
    rv = _gpu(base)
    half = _lib.BtRays()
    half.flux = ctypes.c_void_p(rv.torch("flux").data_ptr())
    ct = bt.CoordTransform(rv.coordSys, rv.coordSys)
    with pytest.raises(RuntimeError):
        _lib.call("bt_rsplit_live", surf._dev.ptr(), _lib.dvec(ct.dr),
                  _lib.dvec(ct.drot.ravel()), med._dev.ptr(), ctypes.byref(coat._record), None,
                  0, rv._rays, half, minFlux, 5, 0, None, _device.stream())
