"""Learned Newton starts: the training kernels against the FP64 numpy
restatement (oracle/mlp.py), the packed MFMA inference against the FP32
inference, the tracer's sampler against the reference intersection, and the
fast-start traces against the reference-start trace (positions and times
within 1e-12 m, directions within 1e-14, flags identical), plus the gain the
network brings over the conic start."""
import ctypes

import numpy as np
import pytest
import torch

import batoid_amd as bt
from batoid_amd import _lib, _device, workloads
from batoid_amd.newton_start import (NewtonStartNet, newton_samples, collect_samples,
                                     train_newton_starts)
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

FIELDS_TRAIN = range(0, 12)
FIELDS_TEST = (100, 150, 188)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _random_problem(n, seed):
    rng = np.random.default_rng(seed)
    net = NewtonStartNet(seed=seed)
    # non-trivial biases and normalisation so every term of the gradient runs
    p = net.params.cpu().numpy().astype(np.float64)
    p[omlp.OFF_B1:omlp.OFF_W2] = rng.normal(0, 0.3, omlp.N_HID)
    p[omlp.OFF_B2:omlp.OFF_W3] = rng.normal(0, 0.3, omlp.N_HID)
    p[omlp.OFF_B3] = 0.1
    net.params = torch.from_numpy(p.astype(np.float32)).cuda()
    feat = rng.normal(0, 1, (n, 6)) * [2.0, 2.0, 0.3, 0.1, 0.1, 1.0] + [0.1, -0.2, 0.05, 0, 0, -1]
    net.in_norm = np.concatenate([[0.1, -0.2, 0.05, 0, 0, -1],
                                  1 / np.array([2.0, 2.0, 0.3, 0.1, 0.1, 1.0])]).astype(np.float32)
    targets = rng.normal(0, 1, n)
    return net, feat.astype(np.float32), targets.astype(np.float32)


def test_loss_grad_matches_oracle():
    net, feat, targ = _random_problem(5000, seed=3)
    p64 = net.params.cpu().numpy().astype(np.float64)
    loss_o, grad_o = omlp.loss_grad(p64, feat.astype(np.float64), targ, net.in_norm)
    f = torch.from_numpy(feat).cuda()
    t = torch.from_numpy(targ).cuda()
    loss, grad = net.loss_grad(f, t)
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    scale = np.abs(grad_o).max()
    assert abs(loss.item() - loss_o) <= 1e-5 * loss_o
    assert np.abs(g - grad_o).max() <= 2e-5 * scale, np.abs(g - grad_o).max() / scale
    # every block of the parameter vector is exercised
    for a, b in ((omlp.OFF_W1, omlp.OFF_B1), (omlp.OFF_B1, omlp.OFF_W2), (omlp.OFF_W2, omlp.OFF_B2),
                 (omlp.OFF_B2, omlp.OFF_W3), (omlp.OFF_W3, omlp.OFF_B3)):
        assert np.abs(grad_o[a:b]).max() > 0


def test_loss_grad_with_index_matches_oracle():
    net, feat, targ = _random_problem(3000, seed=4)
    rng = np.random.default_rng(9)
    idx = rng.integers(0, 3000, 777).astype(np.int32)   # ragged: not a multiple of 64
    p64 = net.params.cpu().numpy().astype(np.float64)
    loss_o, grad_o = omlp.loss_grad(p64, feat[idx].astype(np.float64), targ[idx], net.in_norm)
    loss, grad = net.loss_grad(torch.from_numpy(feat).cuda(), torch.from_numpy(targ).cuda(),
                               torch.from_numpy(idx).cuda())
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - loss_o) <= 1e-5 * loss_o
    assert np.abs(g - grad_o).max() <= 2e-5 * np.abs(grad_o).max()


def test_adam_matches_oracle():
    rng = np.random.default_rng(5)
    n = omlp.N_PARAMS
    # FP32 inputs, so the FP64 oracle sees exactly what the kernel sees
    p, g = rng.normal(0, 1, n).astype(np.float32), rng.normal(0, 1e-2, n).astype(np.float32)
    m = rng.normal(0, 1e-3, n).astype(np.float32)
    v = rng.uniform(0, 1e-4, n).astype(np.float32)
    dev = [torch.from_numpy(a.astype(np.float32)).cuda() for a in (p, g, m, v)]
    _lib.call("bt_mlp_adam", *[_p(t) for t in dev], n, 1e-3, 0.9, 0.999, 1e-8, 7,
              _device.stream())
    torch.cuda.synchronize()
    # hyper-parameters as the kernel holds them (FP32)
    f32 = lambda x: float(np.float32(x))
    po, mo, vo = omlp.adam(*(a.astype(np.float64) for a in (p, g, m, v)), 7, f32(1e-3),
                           beta1=f32(0.9), beta2=f32(0.999), eps=f32(1e-8))
    # FP32 arithmetic: a few units in the last place of each result's scale
    for got, want in ((dev[0], po), (dev[2], mo), (dev[3], vo)):
        ulp = np.spacing(np.float32(np.abs(want).max()))
        assert np.abs(got.cpu().numpy() - want).max() <= 4 * ulp


def test_forward_fp32_and_packed_match_oracle():
    net, feat, _ = _random_problem(10_001, seed=6)     # ragged last wave
    net.out_mean, net.out_std = 3e-7, 2e-7
    y_o = omlp.forward(net.params.cpu().numpy(), feat, net.in_norm) * 2e-7 + 3e-7
    f = torch.from_numpy(feat).cuda()
    y32 = net.predict(f).cpu().numpy()
    y16 = net.predict_packed(f).cpu().numpy().astype(np.float64)
    spread = np.std(y_o)
    assert np.abs(y32 - y_o).max() <= 1e-5 * np.abs(y_o).max()
    # fp16 weights and activations with FP32 accumulation
    err = y16 - y_o
    assert np.sqrt(np.mean(err ** 2)) <= 3e-3 * spread, np.sqrt(np.mean(err ** 2)) / spread
    assert np.abs(err).max() <= 2e-2 * spread


def test_samples_match_reference_intersection():
    optic = workloads.perturbed_tma()
    thx, thy = workloads.tma_fields()
    rv = workloads.tma_rays(optic, thx[3], thy[3], 620e-9, nrad=30, naz=300)
    m1 = optic["M1"]
    feat, dtc, dte, status = newton_samples(m1.surface, m1.coordSys, rv)
    ref = rv.copy()
    bt.intersect(m1.surface, ref, m1.coordSys)
    torch.cuda.synchronize()
    st = status.cpu().numpy()
    ok = st == 3
    assert ok.mean() > 0.99
    assert np.array_equal(ref.failed, (st & 2) == 0)
    # the sampler's dt* reproduces the reference intersection point exactly
    loc = rv.copy().toCoordSys(m1.coordSys)
    x = loc.x + loc.vx * dte.cpu().numpy()
    np.testing.assert_array_equal(x[ok], ref.x[ok])
    # dt_conic is a root of the base conic (closed form of the conic's sag)
    t = dtc.cpu().numpy()
    asph = m1.surface.surfaces[0]
    px, py, pz = loc.x + loc.vx * t, loc.y + loc.vy * t, loc.z + loc.vz * t
    r2 = px * px + py * py
    conic = r2 / (asph.R * (1 + np.sqrt(1 - (1 + asph.conic) * r2 / asph.R ** 2)))
    assert np.abs(pz - conic)[ok].max() < 1e-12
    # and the residual dt* - dt_conic is the micron-scale aspheric term
    assert 1e-9 < np.abs(dte.cpu().numpy() - t)[ok].max() < 1e-4
    # the features are the conic point and the direction (FP32)
    f = feat.cpu().numpy().astype(np.float64)
    for k, c in enumerate((px, py, pz, loc.vx, loc.vy, loc.vz)):
        np.testing.assert_allclose(f[ok, k], c[ok], rtol=2e-7, atol=1e-7)


OPTICS = {"perturbed": "perturbed_tma", "freeform": "freeform_tma"}


@pytest.fixture(scope="module", params=sorted(OPTICS))
def trained(request):
    """Networks for every Newton surface of the perturbed TMA (conic base
    start) and of its freeform variant (polynomial mirrors: plain start)."""
    print(f"case train {request.param}", flush=True)
    optic = getattr(workloads, OPTICS[request.param])()
    thx, thy = workloads.tma_fields()
    rvs = [workloads.tma_rays(optic, thx[k], thy[k], 620e-9, nrad=40, naz=400)
           for k in FIELDS_TRAIN]
    nets = train_newton_starts(optic, rvs, steps=2500, batch=4096)
    return optic, nets


def test_training_converges_and_predicts_residual(trained):
    optic, nets = trained
    # the three perturbed mirrors and the aspheric exit face of L2
    assert sorted(nets) == ["L2_exit", "M1", "M2", "M3"]
    thx, thy = workloads.tma_fields()
    rvs = [workloads.tma_rays(optic, thx[k], thy[k], 620e-9, nrad=30, naz=300)
           for k in FIELDS_TEST]
    held = collect_samples(optic, rvs)
    for name, net in nets.items():
        hist = [l for _, l in net.history]
        assert hist[-1] < 0.05 * hist[0], (name, hist)
        feat, delta = held[name]
        d = delta.cpu().numpy()
        pred = net.predict_packed(feat).cpu().numpy().astype(np.float64)
        resid = pred - d
        gain = np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(resid ** 2))
        print(f"{name}: start error rms {np.sqrt(np.mean(d ** 2)):.3e} m (base) -> "
              f"{np.sqrt(np.mean(resid ** 2)):.3e} m (learned), gain {gain:.1f}", flush=True)
        # the learned start is >= 10x closer to the root than the base one
        assert gain >= 10, (name, gain)


def _cols(rv):
    torch.cuda.synchronize()
    return {k: getattr(rv, k) for k in ("x", "y", "z", "vx", "vy", "vz", "t",
                                        "vignetted", "failed")}


@pytest.mark.parametrize("mode", ["conic", "mlp"])
def test_fast_start_trace_matches_reference(trained, mode):
    optic, nets = trained
    thx, thy = workloads.tma_fields()
    start = nets if mode == "mlp" else "conic"
    for k in FIELDS_TEST:
        for wl in (350e-9, 1050e-9):
            rv = workloads.tma_rays(optic, thx[k], thy[k], wl, nrad=50, naz=500)
            ref = rv.copy()
            optic.trace(ref)
            fast = rv.copy()
            optic.trace(fast, start=start)
            a, b = _cols(ref), _cols(fast)
            assert np.array_equal(a["failed"], b["failed"])
            assert np.array_equal(a["vignetted"], b["vignetted"])
            ok = ~a["failed"]
            assert ok.mean() > 0.99
            for c in ("x", "y", "z", "t"):
                assert np.abs(a[c][ok] - b[c][ok]).max() <= 1e-12, (mode, k, wl, c)
            for c in ("vx", "vy", "vz"):
                assert np.abs(a[c][ok] - b[c][ok]).max() <= 1e-14, (mode, k, wl, c)


def test_fast_start_ragged_sizes_and_partial_networks(trained):
    """Ragged sizes (lanes past the end in the last wave) and a plan whose
    networks cover only some surfaces (the others run the conic start)."""
    optic, nets = trained
    thx, thy = workloads.tma_fields()
    base = workloads.tma_rays(optic, thx[50], thy[50], 620e-9, nrad=13, naz=101)
    for n in (1, 63, 65, 1000, len(base)):
        rv = base[:n] if n < len(base) else base
        ref = rv.copy()
        optic.trace(ref)
        fast = rv.copy()
        optic.trace(fast, start={"M2": nets["M2"]})
        a, b = _cols(ref), _cols(fast)
        assert np.array_equal(a["failed"], b["failed"])
        assert np.array_equal(a["vignetted"], b["vignetted"])
        ok = ~a["failed"]
        for c in ("x", "y", "z", "t"):
            assert np.abs(a[c][ok] - b[c][ok]).max(initial=0) <= 1e-12


def test_conic_start_keeps_the_reference_iteration_budget():
    """Rays whose reference Newton (same start, same iterates) has not
    converged after niter = 5 iterations: the conic start stops there too
    and hands them to the reference path, so their flags are the
    reference's (failed), not intersected by a longer iteration
    (test_oracle_golden.slow_newton_case: 5 iterations fail thousands, 6
    intersect all)."""
    from test_oracle_golden import slow_newton_case
    surf, rays = slow_newton_case()
    optic = bt.CompoundOptic([bt.Interface(surf, name="S")], name="slow",
                             inMedium=bt.vacuum)
    base = bt.RayVector(**{k: rays[k] for k in ("x", "y", "z", "vx", "vy", "vz", "t",
                                                 "wavelength", "flux")},
                        vignetted=rays["vignetted"], failed=rays["failed"])
    ref = base.copy()
    optic.trace(ref)
    fast = base.copy()
    optic.trace(fast, start="conic")
    torch.cuda.synchronize()
    assert ref.failed.sum() > 500
    np.testing.assert_array_equal(fast.failed, ref.failed)
    np.testing.assert_array_equal(fast.vignetted, ref.vignetted)
    ok = ~ref.failed
    for k in ("x", "y", "z", "t"):
        np.testing.assert_allclose(getattr(fast, k)[ok], getattr(ref, k)[ok], rtol=0,
                                   atol=1e-12)


def test_conic_start_on_a_sum_runs_the_reference_iterates():
    """A CONIC start on a Sum not led by an Asphere is the reference's own
    start (the first child's root), so the fast iteration runs the
    reference's iterates and stops within its niter = 5: on rays that need
    more (test_oracle_golden.slow_sum_case: thousands fail at 5, not at 8)
    the flags are the reference's."""
    from test_oracle_golden import slow_sum_case
    surf, rays = slow_sum_case()
    optic = bt.CompoundOptic([bt.Interface(surf, name="S")], name="slow_sum",
                             inMedium=bt.vacuum)
    base = bt.RayVector(**{k: rays[k] for k in ("x", "y", "z", "vx", "vy", "vz", "t",
                                                 "wavelength", "flux")},
                        vignetted=rays["vignetted"], failed=rays["failed"])
    ref = base.copy()
    optic.trace(ref)
    fast = base.copy()
    optic.trace(fast, start="conic")
    torch.cuda.synchronize()
    assert ref.failed.sum() > 5000
    np.testing.assert_array_equal(fast.failed, ref.failed)
    np.testing.assert_array_equal(fast.vignetted, ref.vignetted)
    ok = ~ref.failed
    for k in ("x", "y", "z", "t"):
        np.testing.assert_allclose(getattr(fast, k)[ok], getattr(ref, k)[ok], rtol=0,
                                   atol=1e-12)


@pytest.mark.parametrize("name", ["sum", "sum_plane", "sum_nested", "sum_deep", "sum_wide"])
def test_conic_start_on_every_sum_layout(name):
    """CONIC starts on Sums traced by each plan kernel: Sum(Asphere,
    Zernike, Bicubic) (its Asphere's updated quadric root), a Sum led by a
    Plane, Sums of Sums to three levels and an eleven-child Sum led by a
    Quadric (the first child's reference root).  Flags identical to the
    reference start, positions and times within 1e-12 m, directions within
    1e-14."""
    from helpers import surfaces, random_rays
    surf = dict(surfaces())[name]
    print(f"case conic start {name}", flush=True)
    optic = bt.CompoundOptic([bt.Interface(surf, name="S")], name=name, inMedium=bt.vacuum)
    rays = random_rays(20000, seed=71, z0=1.0, spread=0.15)
    base = bt.RayVector(**{k: rays[k] for k in ("x", "y", "z", "vx", "vy", "vz", "t",
                                                 "wavelength", "flux")},
                        vignetted=rays["vignetted"], failed=rays["failed"])
    ref = base.copy()
    optic.trace(ref)
    fast = base.copy()
    optic.trace(fast, start="conic")
    a, b = _cols(ref), _cols(fast)
    assert np.array_equal(a["failed"], b["failed"])
    assert np.array_equal(a["vignetted"], b["vignetted"])
    ok = ~a["failed"]
    assert ok.mean() > 0.9
    for c in ("x", "y", "z", "t"):
        assert np.abs(a[c][ok] - b[c][ok]).max() <= 1e-12, (name, c)
    for c in ("vx", "vy", "vz"):
        assert np.abs(a[c][ok] - b[c][ok]).max() <= 1e-14, (name, c)
