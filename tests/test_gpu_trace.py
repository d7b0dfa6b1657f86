"""GPU parity: every trace kernel of libbatoid_amd.so against the oracle on
the same seeded inputs.  The trace path is FP64 with no contraction on
either side, so the bar is bit-for-bit identity of every column, NaNs in
the same places and flags equal."""
import ctypes

import numpy as np
import pytest
import torch

import batoid_amd as bt
import oracle
from oracle.host import HostRays, HostPlan, host_records, nthreads
from batoid_amd import _lib
from batoid_amd.plan import PlanSpec, collect_steps
from helpers import (surfaces, media, obscurations, random_rays, transforms,
                     assert_cols_identical, zernike_surface, bicubic_surface, FCOLS, BCOLS)

pytestmark = pytest.mark.gpu

N = 20000
NITER = 5


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


def _gpu_rays(h):
    return bt.RayVector(**{k: h[k] for k in FCOLS}, vignetted=h["vignetted"].astype(bool),
                        failed=h["failed"].astype(bool))


def _host(rv):
    torch.cuda.synchronize()
    return {k: getattr(rv, k) for k in FCOLS + BCOLS}


def _frame(dr, drot):
    """A CoordSys whose transform from the global frame is (dr, drot)."""
    return bt.CoordSys(origin=dr, rot=drot)


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
@pytest.mark.parametrize("action", ["intersect", "reflect", "refract"])
def test_surface_kernels_bit_exact(name, surf, action):
    print(f"case {name}/{action}", flush=True)
    O = oracle.load()
    tab = host_records(surf)
    silica = media()[2][1]
    mtab = host_records(silica)
    for tname, (dr, drot) in transforms():
        for seed, upward in ((1, 0.0), (2, 0.3)):
            base = HostRays(**random_rays(N, seed=seed, upward=upward, spread=0.3))
            ref = base.copy()
            args = (tab.ptr(), _dv(dr), _dv(drot))
            if action == "refract":
                O.or_refract(*args, mtab.ptr(), ref.struct, NITER, nthreads())
            else:
                getattr(O, f"or_{action}")(*args, ref.struct, NITER, nthreads())
            rv = _gpu_rays(base)
            cs = _frame(dr, drot)
            if action == "refract":
                bt.refract(surf, rv, bt.vacuum, silica, coordSys=cs)
            else:
                getattr(bt, action)(surf, rv, coordSys=cs)
            assert_cols_identical(_host(rv), ref.cols, what=f"{name}/{action}/{tname}/{seed}")


@pytest.mark.parametrize("mname,med", media(), ids=[m[0] for m in media()])
def test_refract_media_bit_exact(mname, med):
    O = oracle.load()
    surf = bt.Sphere(2.3)
    tab, mtab = host_records(surf), host_records(med)
    base = HostRays(**random_rays(N, seed=9, n_index=1.0, spread=0.4))
    ref = base.copy()
    O.or_refract(tab.ptr(), _dv(np.zeros(3)), _dv(np.eye(3)), mtab.ptr(), ref.struct,
                 NITER, nthreads())
    rv = _gpu_rays(base)
    bt.refract(surf, rv, bt.vacuum, med)
    assert_cols_identical(_host(rv), ref.cols, what=mname)


def test_total_internal_reflection_bit_exact():
    """Rays leaving glass at grazing angles: NaN velocities and both flags
    (batoid.cpp:478-487)."""
    O = oracle.load()
    surf = bt.Plane()
    med = bt.ConstMedium(1.0)
    base = HostRays(**random_rays(N, seed=10, n_index=1.5, spread=1.0))
    ref = base.copy()
    tab, mtab = host_records(surf), host_records(med)
    O.or_refract(tab.ptr(), _dv(np.zeros(3)), _dv(np.eye(3)), mtab.ptr(), ref.struct,
                 NITER, nthreads())
    tir = np.isnan(ref["vx"]).sum()
    assert tir > 100
    rv = _gpu_rays(base)
    bt.refract(surf, rv, bt.ConstMedium(1.5), med)
    assert_cols_identical(_host(rv), ref.cols, what="TIR")


@pytest.mark.parametrize("oname,obsc", obscurations(), ids=[o[0] for o in obscurations()])
def test_obscurations_bit_exact(oname, obsc):
    O = oracle.load()
    tab = host_records(obsc)
    base = HostRays(**random_rays(N, seed=12))
    ref = base.copy()
    O.or_obscure(tab.ptr(), tab.count, ref.struct, nthreads())
    rv = _gpu_rays(base)
    obsc.obscure(rv)
    assert_cols_identical(_host(rv), ref.cols, what=oname)
    rng = np.random.default_rng(13)
    x, y = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    oc = np.empty(N, np.uint8)
    O.or_contains_n(tab.ptr(), tab.count, x.ctypes.data, y.ctypes.data, oc.ctypes.data, N,
                    nthreads())
    np.testing.assert_array_equal(obsc.contains(x, y), oc.astype(bool))


def test_circular_edges_bit_exact():
    """Points on the edge of circular obscurations (a hexapolar ring at the
    obscuration radius, as a pupil's outermost ring meets a mirror's clear
    aperture): contains() is decided by the last bit of hypot, and the
    GPU's ref_hypot reproduces libm's, so the flags match the oracle's
    (which calls libm) point for point."""
    O = oracle.load()
    th = np.arange(200_000) * (2 * np.pi / 200_000)
    for obsc, r in [(bt.ObscCircle(4.1), 4.1), (bt.ObscNegation(bt.ObscCircle(4.1)), 4.1),
                    (bt.ObscAnnulus(2.558, 4.18), 4.18), (bt.ObscAnnulus(2.558, 4.18), 2.558),
                    (bt.ObscCircle(0.5, 0.1, -0.2), 0.5)]:
        x0, y0 = (0.1, -0.2) if r == 0.5 else (0.0, 0.0)
        x, y = x0 + r * np.cos(th), y0 + r * np.sin(th)
        tab = host_records(obsc)
        oc = np.empty(len(x), np.uint8)
        O.or_contains_n(tab.ptr(), tab.count, x.ctypes.data, y.ctypes.data, oc.ctypes.data,
                        len(x), nthreads())
        got = obsc.contains(x, y)
        assert 0 < oc.sum() < len(x)       # the ring really straddles the edge
        np.testing.assert_array_equal(got, oc.astype(bool), err_msg=repr(obsc))


@pytest.mark.parametrize("name,surf", surfaces(), ids=[s[0] for s in surfaces()])
def test_sag_normal_bit_exact(name, surf):
    O = oracle.load()
    tab = host_records(surf)
    rng = np.random.default_rng(14)
    x, y = rng.uniform(-1.5, 1.5, N), rng.uniform(-1.5, 1.5, N)
    s = np.empty(N)
    nrm = np.empty(3 * N)
    O.or_sag_n(tab.ptr(), x.ctypes.data, y.ctypes.data, s.ctypes.data, N, nthreads())
    O.or_normal_n(tab.ptr(), x.ctypes.data, y.ctypes.data, nrm.ctypes.data, N, nthreads())
    assert_cols_identical({"s": surf.sag(x, y)}, {"s": s}, cols=("s",), what=name)
    g = surf.normal(x, y)
    assert_cols_identical({"n": g.T.ravel()}, {"n": nrm}, cols=("n",), what=name)


@pytest.mark.parametrize("mname,med", media(), ids=[m[0] for m in media()])
def test_media_bit_exact(mname, med):
    O = oracle.load()
    tab = host_records(med)
    w = np.random.default_rng(15).uniform(250e-9, 1200e-9, N)
    o = np.empty(N)
    O.or_get_n_n(tab.ptr(), w.ctypes.data, o.ctypes.data, N, nthreads())
    assert_cols_identical({"n": med.getN(w)}, {"n": o}, cols=("n",), what=mname)


def test_transforms_and_propagation_bit_exact():
    O = oracle.load()
    for tname, (dr, drot) in transforms():
        base = HostRays(**random_rays(N, seed=16))
        ref = base.copy()
        O.or_apply_forward_transform(_dv(dr), _dv(drot), ref.struct, nthreads())
        rv = _gpu_rays(base)
        rv.toCoordSys(_frame(dr, drot))
        assert_cols_identical(_host(rv), ref.cols, what=f"forward {tname}")
        O.or_apply_reverse_transform(_dv(dr), _dv(drot), ref.struct, nthreads())
        bt.CoordTransform(bt.CoordSys(), _frame(dr, drot)).applyReverse(rv)
        assert_cols_identical(_host(rv), ref.cols, what=f"reverse {tname}")
    base = HostRays(**random_rays(N, seed=17))
    ref = base.copy()
    O.or_propagate_in_place(ref.struct, 3.25, nthreads())
    rv = _gpu_rays(base)
    pos = rv.positionAtTime(3.25)
    rv.propagate(3.25)
    assert_cols_identical(_host(rv), ref.cols, what="propagate")
    assert_cols_identical({"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2]}, ref.cols,
                          cols=("x", "y", "z"), what="positionAtTime")


def test_field_operations():
    O = oracle.load()
    base = HostRays(**random_rays(N, seed=18))
    rv = _gpu_rays(base)
    p = np.empty(N)
    O.or_phase(base.struct, 0.1, -0.2, 0.3, 1.5, p.ctypes.data, nthreads())
    assert_cols_identical({"p": rv.phase([0.1, -0.2, 0.3], 1.5)}, {"p": p}, cols=("p",))
    a = np.empty(2 * N)
    O.or_amplitude(base.struct, 0.1, -0.2, 0.3, 1.5, a.ctypes.data, nthreads())
    # exp(i phi): FP64 sin/cos of the device math library vs libm, <= 1 ulp
    np.testing.assert_allclose(rv.amplitude([0.1, -0.2, 0.3], 1.5).view(np.float64), a,
                               rtol=0, atol=4e-16 * np.max(np.abs(p)))
    for ign in (True, False):
        s = np.empty(2)
        O.or_sum_amplitude(base.struct, 0.1, -0.2, 0.3, 1.5, int(ign), s.ctypes.data, 1)
        g = rv.sumAmplitude([0.1, -0.2, 0.3], 1.5, ignoreVignetted=ign)
        assert abs(g - complex(s[0], s[1])) <= 1e-12 * N


def test_fused_plan_matches_per_surface_and_oracle():
    """Whole-optic fused kernel == oracle's per-Interface passes, on an optic
    that mixes every interaction, media, obscurations and a perturbed Sum."""
    from helpers import zernike_surface, bicubic_surface
    silica = bt.SellmeierMedium(0.6961663, 0.4079426, 0.8974794, 0.00467914825849,
                                0.013512063073959999, 97.93400253792099)
    cs = bt.CoordSys()
    m1 = bt.Mirror(bt.Sum(bt.Asphere(19.835, -1.215, [0.0, -1.381e-9]),
                          zernike_surface(jmax=15, R_outer=1.0), bicubic_surface()),
                   name="M1", coordSys=cs, obscuration=bt.ObscNegation(bt.ObscAnnulus(0.1, 1.0)))
    l1 = bt.Lens([
        bt.RefractiveInterface(bt.Sphere(-2.0), name="L1a", coordSys=cs.shiftLocal([0, 0, 5.0]),
                               inMedium=bt.vacuum, outMedium=silica,
                               obscuration=bt.ObscNegation(bt.ObscCircle(0.7))),
        bt.RefractiveInterface(bt.Quadric(-3.0, -0.5), name="L1b",
                               coordSys=cs.shiftLocal([0.001, 0, 5.05]).rotateLocal(bt.RotX(0.002)),
                               inMedium=silica, outMedium=bt.vacuum),
    ], name="L1", inMedium=bt.vacuum, outMedium=bt.vacuum)
    baffle = bt.Baffle(bt.Plane(), name="B", coordSys=cs.shiftLocal([0, 0, 6.0]),
                       obscuration=bt.ObscRectangle(0.05, 3.0))
    det = bt.Detector(bt.Tilted(0.001, 0.0), name="D", coordSys=cs.shiftLocal([0, 0, 9.0]))
    optic = bt.CompoundOptic([m1, l1, baffle, det], name="test", inMedium=bt.vacuum)

    base = HostRays(**random_rays(N, seed=19, z0=2.0, spread=0.02, radius=1.0))
    spec = PlanSpec(collect_steps(optic), bt.CoordSys())
    ref = HostPlan(spec).trace(base.copy(), niter=NITER)
    assert 0 < ref["vignetted"].sum() < N

    rv = _gpu_rays(base)
    optic.trace(rv)
    assert rv.coordSys == det.coordSys
    assert_cols_identical(_host(rv), ref.cols, what="fused plan")

    rv2 = _gpu_rays(base)
    for item in [m1, l1.items[0], l1.items[1], baffle, det]:
        item.trace(rv2)
    assert_cols_identical(_host(rv2), ref.cols, what="per-surface")


def test_perturbed_tma_bit_exact():
    """C3 optic: every mirror a Sum(Asphere, Zernike j<=28, Bicubic 204^2);
    the fused trace equals the oracle bit for bit."""
    from batoid_amd import workloads
    optic = workloads.perturbed_tma()
    thx, thy = workloads.tma_fields()
    for k, wl in ((0, 350e-9), (7, 1050e-9)):
        rv = workloads.tma_rays(optic, thx[k], thy[k], wl, nrad=60, naz=600)
        ref = HostRays.from_rv(rv)
        HostPlan(PlanSpec(collect_steps(optic), rv.coordSys)).trace(ref, niter=NITER)
        optic.trace(rv)
        got = _host(rv)
        assert_cols_identical(got, ref.cols, what=f"perturbed TMA field {k}")
        assert 0 < got["vignetted"].sum() < len(rv)


def test_plan_kernels_agree_on_sum_layouts():
    """The perturbed TMA's plan runs the kernel with its Sum(Asphere,
    Polynomial, Bicubic) mirrors unrolled; the all-kinds kernel on the same
    plan gives the same bits.  The freeform TMA (Sum(Polynomial, Bicubic))
    runs the all-kinds kernel, bit-identical to the oracle."""
    from batoid_amd import workloads
    from batoid_amd import _device
    optic = workloads.perturbed_tma()
    thx, thy = workloads.tma_fields()
    rv = workloads.tma_rays(optic, thx[3], thy[3], 770e-9, nrad=60, naz=600)
    plan = optic._plan(rv.coordSys, False)
    assert plan.struct.flags & _lib.PLAN_SUM_APB
    a = rv.copy()
    optic.trace(a)
    b = rv.copy()
    for other in (0, _lib.PLAN_FLAT_SUMS):            # nested-capable, flat Sums
        b = rv.copy()
        st = _lib.BtPlan.from_buffer_copy(plan.struct)
        st.flags = (st.flags & ~_lib.PLAN_SUM_APB) | other
        _lib.call("bt_trace_plan", st, b._rays, b._rays, NITER, _device.stream())
        assert_cols_identical(_host(b), _host(a), what=f"kernel {other} vs APB kernel")

    free = workloads.freeform_tma()
    rv = workloads.tma_rays(free, thx[3], thy[3], 770e-9, nrad=60, naz=600)
    fplan = free._plan(rv.coordSys, False)
    assert fplan.struct.flags & _lib.PLAN_FLAT_SUMS
    assert not fplan.struct.flags & _lib.PLAN_SUM_APB
    ref = HostRays.from_rv(rv)
    HostPlan(PlanSpec(collect_steps(free), rv.coordSys)).trace(ref, niter=NITER)
    b = rv.copy()
    free.trace(rv)
    assert_cols_identical(_host(rv), ref.cols, what="freeform TMA")
    st = _lib.BtPlan.from_buffer_copy(fplan.struct)
    st.flags &= ~_lib.PLAN_FLAT_SUMS
    _lib.call("bt_trace_plan", st, b._rays, b._rays, NITER, _device.stream())
    assert_cols_identical(_host(b), ref.cols, what="freeform TMA, nested-capable kernel")


def test_nested_sum_plan_bit_exact():
    """Mirrors whose surfaces are Sums of Sums (three levels) and a Sum of
    eleven children, traced by the fused plan (the nested-capable kernel):
    bit-identical to the oracle, which equals the reference's compiled core
    on these surfaces (test_oracle_vs_ref).  With the flat-Sum flag forced
    on, the nested records are not read as leaves' children: such rays fail."""
    from batoid_amd import _device
    s = dict(surfaces())
    cs = bt.CoordSys()
    for name in ("sum_nested", "sum_deep", "sum_wide"):
        optic = bt.CompoundOptic([
            bt.Mirror(s[name], name="M", coordSys=cs,
                      obscuration=bt.ObscNegation(bt.ObscCircle(1.0))),
            bt.Detector(bt.Plane(), name="D", coordSys=cs.shiftLocal([0, 0, 2.5]))],
            name=name, inMedium=bt.vacuum)
        base = HostRays(**random_rays(N, seed=31, z0=2.0, spread=0.05, radius=1.1))
        spec = PlanSpec(collect_steps(optic), bt.CoordSys())
        nested = name != "sum_wide"
        assert bool(spec.flags & _lib.PLAN_FLAT_SUMS) == (not nested)
        ref = HostPlan(spec).trace(base.copy(), niter=NITER)
        assert 0 < ref["vignetted"].sum() < N
        rv = _gpu_rays(base)
        optic.trace(rv)
        assert_cols_identical(_host(rv), ref.cols, what=f"plan {name}")
        if nested:
            rv = _gpu_rays(base)
            st = _lib.BtPlan.from_buffer_copy(optic._plan(rv.coordSys, False).struct)
            st.flags |= _lib.PLAN_FLAT_SUMS
            _lib.call("bt_trace_plan", st, rv._rays, rv._rays, NITER, _device.stream())
            assert rv.failed[~base["failed"].astype(bool)].all()


def test_sum_layout_flag_on_other_sums_fails_rays():
    """A plan flagged BT_PLAN_SUM_APB whose Sum has another layout fails the
    rays that reach it (the kernel checks the layout before reading the
    children), rather than reading the wrong records."""
    from batoid_amd import _device
    optic = bt.CompoundOptic([
        bt.Mirror(bt.Sum(bt.Plane(), bt.Paraboloid(12.0)), name="M"),
        bt.Detector(bt.Plane(), name="D", coordSys=bt.CoordSys(origin=[0, 0, 3.0]))],
        name="sum2", inMedium=bt.vacuum)
    base = HostRays(**random_rays(4096, seed=23, z0=2.0, spread=0.02, prefail=0.0))
    rv = _gpu_rays(base)
    plan = optic._plan(rv.coordSys, False)
    assert not plan.struct.flags & (_lib.PLAN_SUM_APB | _lib.PLAN_ANALYTIC)
    st = _lib.BtPlan.from_buffer_copy(plan.struct)
    st.flags |= _lib.PLAN_SUM_APB
    _lib.call("bt_trace_plan", st, rv._rays, rv._rays, NITER, _device.stream())
    assert rv.failed.all() and rv.vignetted.all()


def _crafted_zernikes():
    """Zernike records whose coefficient tables exercise the trimmed Horner
    loops: the triangular tables of a real series, tables with -0.0 entries
    (kept), rows of only +0.0 inside and at the top, a row of only -0.0, and
    an all-zero table."""
    rng = np.random.default_rng(77)
    out = []
    base = zernike_surface(seed=71, jmax=28)
    out.append(("triangular", base))
    z = zernike_surface(seed=72, jmax=28)
    c, gx, gy = (t.copy() for t in (z._xycoef, z._xycoef_gradx, z._xycoef_grady))
    c[2, :] = 0.0                      # a row of +0.0 inside
    c[6, :] = 0.0                      # the top row: +0.0 only
    c[5, :] = 0.0
    c[4, 3] = -0.0                     # a -0.0 beyond the row's last nonzero
    c[1, 6] = -0.0
    gx[3, :] = -0.0                    # a row of only -0.0
    gy[5, :] = 0.0
    gy[4, :] = 0.0
    gy[0, 5] = rng.normal() * 1e-4     # a nonzero past the triangle
    z._xycoef, z._xycoef_gradx, z._xycoef_grady = c, gx, gy
    out.append(("signed_zeros", z))
    z = zernike_surface(seed=73, jmax=28)
    z._xycoef = np.zeros_like(z._xycoef)
    z._xycoef_gradx = np.zeros_like(z._xycoef_gradx)
    z._xycoef_grady = np.zeros_like(z._xycoef_grady)
    out.append(("all_zero", z))
    return out


def test_polynomial_trim_bit_exact():
    """Polynomial sags and normals skip the +0.0 runs of their coefficient
    tables; the oracle runs the reference's full loops.  Points include
    +-0, NaN, +-inf and huge coordinates, alone in their waves and mixed
    into waves of ordinary points (any non-finite lane sends its wave
    through the full loop)."""
    O = oracle.load()
    rng = np.random.default_rng(78)
    x, y = rng.uniform(-1.5, 1.5, N), rng.uniform(-1.5, 1.5, N)
    special = np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 1e300, -1e-300, 5e-324])
    k = len(special)
    x[:k * k] = np.repeat(special, k)          # every pair, in the first waves
    y[:k * k] = np.tile(special, k)
    idx = rng.choice(np.arange(256, N), 64, replace=False)
    x[idx] = special[rng.integers(0, k, 64)]   # stray ones in ordinary waves
    y[idx[::2]] = special[rng.integers(0, k, 32)]
    for name, surf in _crafted_zernikes():
        trim = surf._leaf_record()[1][3]
        assert trim is not None and trim.dtype == np.int32
        tab = host_records(surf)
        s = np.empty(N)
        nrm = np.empty(3 * N)
        with np.errstate(all="ignore"):
            O.or_sag_n(tab.ptr(), x.ctypes.data, y.ctypes.data, s.ctypes.data, N, nthreads())
            O.or_normal_n(tab.ptr(), x.ctypes.data, y.ctypes.data, nrm.ctypes.data, N,
                          nthreads())
        assert_cols_identical({"s": surf.sag(x, y)}, {"s": s}, cols=("s",), what=name)
        g = surf.normal(x, y)
        assert_cols_identical({"n": g.T.ravel()}, {"n": nrm}, cols=("n",), what=name)
        # the same surfaces inside a Sum, traced by the plan kernel
        optic = bt.CompoundOptic([
            bt.Mirror(bt.Sum(bt.Asphere(9.0, -0.9, [1e-4]), surf, bicubic_surface(seed=9)),
                      name="M", coordSys=bt.CoordSys()),
            bt.Detector(bt.Plane(), name="D", coordSys=bt.CoordSys().shiftLocal([0, 0, 2.5]))],
            name=name, inMedium=bt.vacuum)
        base = HostRays(**random_rays(N, seed=79, z0=2.0, spread=0.05, radius=1.1))
        ref = HostPlan(PlanSpec(collect_steps(optic), bt.CoordSys())).trace(base.copy(),
                                                                             niter=NITER)
        rv = _gpu_rays(base)
        optic.trace(rv)
        assert_cols_identical(_host(rv), ref.cols, what=f"plan {name}")
