"""Batched (field, wavelength) PSFs: one fused trace + one persistent MFMA
GEMM launch over uneven segments, against the oracle's FP64 direct sum per
segment (relative L2 <= 1e-5), plus graph capture and replay (every replay
bit-identical to the eager run, new input data picked up), segments with no
summable rays, the 'mean' reference point and the autotuner."""
import numpy as np
import pytest
import torch

import batoid_amd as bt
import oracle
from oracle.host import HostRays, nthreads
from batoid_amd import workloads
from batoid_amd.analysis import _chief_index
from batoid_amd.batch import PSFBatch

pytestmark = pytest.mark.gpu

NX = [24, 31, 40, 17, 32, 28]          # uneven segments, odd and even grids
NPIX = 32
DX = 0.5e-6


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def _oracle_direct(rays, point, prim, n, time):
    h = HostRays.from_rv(rays)
    out = np.empty(2 * n * n)
    oracle.load().or_huygens_direct(
        h.struct, bt._lib.dvec(point[:2]), bt._lib.dvec(prim[0]), bt._lib.dvec(prim[1]),
        n, n, float(time), 1, out.ctypes.data, nthreads())
    return out.view(np.complex128).reshape(n, n)


def _batch(reference='chief', precision='fp32', **kw):
    optic = workloads.tma()
    thx, thy = workloads.tma_fields(n=3, seed=1)
    fx = np.repeat(thx, 2)
    fy = np.repeat(thy, 2)
    wl = np.tile([490e-9, 910e-9], 3)
    return PSFBatch(optic, fx, fy, wl, NX, nxOut=NPIX, dx=DX, reference=reference,
                    precision=precision, **kw)


def _segment(b, s):
    return b.rays[int(b.bounds[s]):int(b.bounds[s + 1])]


def _check_against_oracle(b, reference, tol=1e-5):
    amp = b.amplitudes()
    for s in range(b.nseg):
        rays = _segment(b, s)
        if reference == 'chief':
            point = rays.r[_chief_index(NX[s])]
        else:
            point = np.mean(rays.r[~rays.vignetted], axis=0)
        ref = _oracle_direct(rays, point, b.lattices[s], NPIX, float(rays.t[0]))
        err = rel_l2(amp[s], ref)
        print(f"segment {s} nx={NX[s]}: rel L2 {err:.2e}", flush=True)
        assert err < tol, (s, err)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_batch_matches_oracle_per_segment(precision):
    print(f"case batch chief {precision}", flush=True)
    b = _batch(precision=precision, chunk=64, nwg=37)   # many items, odd grid
    b.run()
    _check_against_oracle(b, 'chief')


def test_batch_mean_reference():
    print("case batch mean", flush=True)
    b = _batch(reference='mean', chunk=256)
    b.run()
    _check_against_oracle(b, 'mean')


def test_batch_trace_matches_single_traces():
    """The batched trace is the per-segment trace, bit for bit."""
    b = _batch()
    b.run()
    torch.cuda.synchronize()
    for s in range(b.nseg):
        single = b.rays_in[int(b.bounds[s]):int(b.bounds[s + 1])]
        b.optic.trace(single)
        got = _segment(b, s)
        for c in ("x", "y", "z", "vx", "vy", "vz", "t", "vignetted", "failed"):
            np.testing.assert_array_equal(getattr(got, c), getattr(single, c))


def test_segment_without_rays_is_zero():
    b = _batch(chunk=128)
    lo, hi = int(b.bounds[2]), int(b.bounds[3])
    b.rays_in.torch("failed")[lo:hi] = 1
    b.run()
    amp = b.amplitudes()
    assert np.all(amp[2] == 0)
    assert np.abs(amp[1]).max() > 0
    assert b.valid_rays()[2] == 0


def test_graph_replays_match_eager():
    b = _batch(chunk=512)
    b.run()
    eager = b.amplitudes().copy()
    eager_x = b.rays.x.copy()
    b.capture()
    for k in range(3):
        b.amp.zero_()
        # unrelated eager work between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        b.run()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(b.amplitudes(), eager)
        np.testing.assert_array_equal(b.rays.x, eager_x)
        del junk
    # new input data in the same buffers is picked up by the replay
    b.rays_in.torch("wavelength").mul_(1.25)
    b.run()
    replay = b.amplitudes().copy()
    b.graph = None
    b.run()
    np.testing.assert_array_equal(replay, b.amplitudes())
    assert rel_l2(replay, eager) > 1e-3


def test_autotune_keeps_results():
    b = _batch()
    b.run()
    first = b.amplitudes().copy()
    timings = b.autotune(chunks=(64, 1024), nwgs=(64, 512), reps=1)
    assert len(timings) == 4 and all(t > 0 for t in timings.values())
    b.run()
    # only the FP32 summation order inside work items changes
    assert rel_l2(b.amplitudes(), first) < 1e-5
