"""Batched (field, wavelength) traces and Huygens PSFs in one launch sequence.

The reference computes one PSF per ``huygensPSF`` call (analysis.py:181-305):
generate a ray grid, trace it, and sum amplitudes pixel by pixel.  A survey
of many fields and wavelengths is a Python loop over such calls; the nearest
batched form in the reference is concatenate-then-trace
(rayVector.py:1351-1365, analysis.py:116-124).

``PSFBatch`` lays all segments out as one ray batch in HBM and runs

    bt_trace_plan      -- the fused whole-optic trace over every segment
                          (each ray carries its own wavelength)
    bt_huygens_batch   -- per-segment reference points and times, one
                          compaction, and one persistent MFMA GEMM launch
                          over (segment, ray chunk) work items

on one stream.  Nothing in that sequence synchronises with the host or
allocates, so it is captured once as a HIP graph (``capture``) and replayed
(``run``); new ray data written into ``rays_in`` is picked up by the replay.
``autotune`` times candidate (chunk, workgroup) settings on the batch itself.
"""
import numpy as np
import torch

from . import _lib, _device
from .analysis import _chief_index, _reciprocalLatticeVectors, dkdu, _PRECISION
from .lattice import Lattice
from .rayVector import RayVector, concatenateRayVectors, _COLS
from .utils import fieldToDirCos


def _per_segment(v, nseg, dtype):
    a = np.asarray(v, dtype=dtype)
    if a.ndim == 0:
        return np.full(nseg, a.item(), dtype=dtype)
    if a.shape != (nseg,):
        raise ValueError(f"expected a scalar or {nseg} values, got shape {a.shape}")
    return a


class PSFBatch:
    """PSFs of ``optic`` for segments s = (theta_x[s], theta_y[s],
    wavelength[s]) with an nx[s] x nx[s] pupil grid each, all on one output
    lattice size ``nxOut * pad`` (pad = 2 with the default fftPSF lattice,
    1 when ``dx`` is given), as huygensPSF would compute them one by one.

    ``start`` is passed to the trace (see CompoundOptic.trace)."""

    def __init__(self, optic, theta_x, theta_y, wavelength, nx, nxOut=None, dx=None,
                 dy=None, reference='chief', projection='postel', precision='fp32',
                 start=None, chunk=2048, nwg=0):
        from numbers import Real
        thx = np.atleast_1d(np.asarray(theta_x, dtype=float))
        nseg = len(thx)
        thy = _per_segment(theta_y, nseg, float)
        wl = _per_segment(wavelength, nseg, float)
        nxs = _per_segment(nx, nseg, np.int64)
        if nxOut is None:
            if np.any(nxs != nxs[0]):
                raise ValueError("nxOut is required when segments have different nx")
            nxOut = int(nxs[0])
        if reference not in ('chief', 'mean'):
            raise ValueError(f"unknown reference {reference!r}")
        if precision not in _PRECISION:
            raise ValueError(f"unknown precision {precision!r}")
        self.optic = optic
        self.nseg = nseg
        self.precision = precision
        self.start = start

        # lattices (analysis.py:234-258)
        lattices = []
        if dx is None:
            pad = 2
            for s in range(nseg):
                d = nxs[s] - 2 if (nxs[s] % 2) == 0 else nxs[s] - 1
                primitiveU = np.array([[optic.pupilSize / d, 0], [0, optic.pupilSize / d]])
                primitiveK = dkdu(optic, thx[s], thy[s], wl[s],
                                  projection=projection).dot(primitiveU)
                lattices.append(np.vstack(_reciprocalLatticeVectors(
                    primitiveK[0], primitiveK[1], pad * nxs[s])))
        else:
            pad = 1
            if isinstance(dx, Real):
                prim = np.vstack([[dx, 0], [0, dx if dy is None else dy]])
            else:
                prim = np.vstack([dx, dy])
            lattices = [prim] * nseg
        self.lattices = lattices
        self.npix = int(nxOut * pad)

        # rays: one asGrid per segment (analysis.py:262-265), concatenated
        rvs = []
        for s in range(nseg):
            dirCos = fieldToDirCos(thx[s], thy[s], projection=projection)
            rvs.append(RayVector.asGrid(optic=optic, wavelength=wl[s], dirCos=dirCos,
                                        nx=int(nxs[s])))
        sizes = np.array([len(r) for r in rvs], dtype=np.int64)
        self.bounds = np.concatenate([[0], np.cumsum(sizes)])
        self.rays_in = concatenateRayVectors(rvs)
        del rvs
        n = len(self.rays_in)
        self.n = n
        dev = _device.device()
        # output columns; wavelength and flux are shared with the input
        cols = {k: torch.empty_like(self.rays_in.torch(k))
                for k in _COLS if k not in ("wavelength", "flux")}
        cols["wavelength"] = self.rays_in.torch("wavelength")
        cols["flux"] = self.rays_in.torch("flux")
        self.plan = optic._plan(self.rays_in.coordSys, False, start)
        self.rays = RayVector._directInit(cols, self.plan.spec.coordSys_out)

        segs = (_lib.BtPsfSegment * nseg)()
        for s in range(nseg):
            b, e = int(self.bounds[s]), int(self.bounds[s + 1])
            segs[s].begin, segs[s].end = b, e
            segs[s].chief = b + _chief_index(int(nxs[s])) if reference == 'chief' else -1
            segs[s].time_ray = b
            for k in range(2):
                segs[s].a1[k] = float(lattices[s][0][k])
                segs[s].a2[k] = float(lattices[s][1][k])
        self.segments = _device.upload_bytes(bytes(memoryview(segs).cast("B")))
        self.amp = torch.empty((nseg, self.npix, self.npix, 2), dtype=torch.float64, device=dev)
        self.graph = None
        self._configure(chunk, nwg)

    def set_start(self, start):
        """Change the trace's Newton start (see CompoundOptic.trace); a
        captured graph is dropped."""
        self.start = start
        self.plan = self.optic._plan(self.rays_in.coordSys, False, start)
        self.graph = None

    # ------------------------------------------------------------------
    def _configure(self, chunk, nwg):
        lib = _lib.load()
        nbytes = int(lib.bt_huygens_batch_workspace_bytes(self.n, self.nseg, self.npix,
                                                          self.npix, int(chunk)))
        if nbytes <= 0:
            raise _lib.BatoidAMDError(f"bad batch configuration (chunk={chunk})")
        self.chunk, self.nwg = int(chunk), int(nwg)
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=_device.device())
        self.graph = None

    def _launch(self):
        from .global_vars import _batoid_niter
        stream = _device.stream()
        if self.plan.struct.nstep > 0:
            _lib.call("bt_trace_plan", self.plan.struct, self.rays_in._rays, self.rays._rays,
                      _batoid_niter, stream)
        _lib.call("bt_huygens_batch", self.rays._rays, _device.ptr(self.segments), self.nseg,
                  self.npix, self.npix, 1, _PRECISION[self.precision], self.chunk, self.nwg,
                  _device.ptr(self.amp), _device.ptr(self.workspace), self.workspace.numel(),
                  stream)

    def capture(self):
        """Record the launch sequence as one HIP graph (after an eager run,
        which loads every kernel)."""
        self._launch()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch()
        self.graph = g
        return g

    def run(self):
        """Trace every segment and compute every PSF (graph replay once
        captured).  Asynchronous: results are in ``amp`` / ``rays``."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch()
        return self

    def time(self, reps=5):
        """Mean milliseconds per run()."""
        self.run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            self.run()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def autotune(self, chunks=(1024, 2048, 4096, 8192), nwgs=(256, 512, 768, 1024), reps=3):
        """Pick the (chunk, workgroups) pair with the shortest eager run on
        this batch; returns {(chunk, nwg): ms}.  Results do not depend on the
        choice beyond FP32 summation order within a segment."""
        graph = self.graph is not None
        timings = {}
        for c in chunks:
            for w in nwgs:
                self._configure(c, w)
                timings[(c, w)] = self.time(reps)
        best = min(timings, key=timings.get)
        self._configure(*best)
        if graph:
            self.capture()
        return timings

    # ------------------------------------------------------------------
    def amplitudes(self):
        torch.cuda.synchronize()
        return _device.to_numpy(self.amp).view(np.complex128)[..., 0]

    def psfs(self):
        """One Lattice per segment, |amplitude|^2 (analysis.py:304-305)."""
        amp = self.amplitudes()
        return [Lattice(np.abs(amp[s]) ** 2, self.lattices[s]) for s in range(self.nseg)]

    def valid_rays(self):
        """Rays summed per segment (not vignetted, not failed)."""
        torch.cuda.synchronize()
        ok = ((self.rays.torch("vignetted") == 0) & (self.rays.torch("failed") == 0)).cpu().numpy()
        return np.add.reduceat(ok.astype(np.int64), self.bounds[:-1]) if self.n else \
            np.zeros(self.nseg, dtype=np.int64)


def huygensPSFBatch(optic, theta_x, theta_y, wavelength, nx, nxOut=None, dx=None, dy=None,
                    reference='chief', projection='postel', precision='fp32', start=None):
    """huygensPSF (analysis.py:181-305) for many (field, wavelength) segments
    at once: a list of Lattices, one per segment."""
    b = PSFBatch(optic, theta_x, theta_y, wavelength, nx, nxOut=nxOut, dx=dx, dy=dy,
                 reference=reference, projection=projection, precision=precision, start=start)
    return b.run().psfs()
