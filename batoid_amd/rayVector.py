"""RayVector: structure-of-arrays ray columns resident in GPU memory.

Same constructor, generators and accessors as batoid/rayVector.py.  The
11 columns (x, y, z, vx, vy, vz, t, wavelength, flux as FP64; vignetted,
failed as u8) are torch device tensors; the HIP library receives their raw
pointers.  Column accessors (``rv.x`` ...) return host numpy copies, the
counterpart of the reference's syncToHost on property access
(rayVector.py:1083-1176); ``rv.torch(name)`` gives the device column
itself without a copy.
"""
import ctypes
from numbers import Real

import numpy as np
import torch

from . import _lib, _device, _pcg
from .constants import globalCoordSys, vacuum
from .coordTransform import CoordTransform
from .utils import fieldToDirCos

_FCOLS = ("x", "y", "z", "vx", "vy", "vz", "t", "wavelength", "flux")
_BCOLS = ("vignetted", "failed")
_COLS = _FCOLS + _BCOLS


def _hexapolar_rings(outer, inner, nrad, naz, kfold=6):
    """Radii, azimuth counts and azimuth steps of the hexapolar rings,
    outermost first (utils.py:512-531): nphi = (naz rho / outer) // kfold *
    kfold (kfold when that is 0; 1 at rho = 0 when inner is 0), with the
    same float64 operations as the reference's per-ring loop, elementwise."""
    rhos = np.linspace(outer, inner, nrad)
    nphis = (np.floor_divide(naz * rhos / outer, kfold)).astype(np.int64) * kfold
    nphis[nphis == 0] = kfold
    if inner == 0.0:
        nphis[-1] = 1
        rhos[-1] = 0.0
    return rhos, nphis, 2 * np.pi / nphis


class RayVector:
    """Rays as device columns.  Arguments broadcast against each other and
    are copied (rayVector.py:21-77)."""

    def __init__(self, x, y, z, vx, vy, vz, t=0.0, wavelength=0.0, flux=1.0,
                 vignetted=False, failed=False, coordSys=globalCoordSys):
        shape = np.broadcast(x, y, z, vx, vy, vz, t, wavelength, flux,
                             vignetted, failed).shape
        if shape == ():
            shape = (1,)
        cols = {}
        for name, val in zip(_FCOLS, (x, y, z, vx, vy, vz, t, wavelength, flux)):
            if isinstance(val, torch.Tensor):
                val = val.detach().cpu().numpy()
            cols[name] = _device.f64(np.broadcast_to(np.asarray(val, float), shape).ravel())
        for name, val in zip(_BCOLS, (vignetted, failed)):
            if isinstance(val, torch.Tensor):
                val = val.detach().cpu().numpy()
            cols[name] = _device.u8(np.broadcast_to(np.asarray(val, bool), shape).ravel())
        self._cols = cols
        self._buf = None
        self._rays_c = None
        self.coordSys = coordSys

    @staticmethod
    def _directInit(cols, coordSys, buf=None):
        """Wrap existing device columns (no copy); ``buf`` is the one buffer
        they are views of, when they are."""
        ret = RayVector.__new__(RayVector)
        ret._cols = cols
        ret._buf = buf
        ret._rays_c = None
        ret.coordSys = coordSys
        return ret

    @staticmethod
    def _layout(n):
        """Byte offsets of the columns of n rays in one allocation (each
        column starting on a 256-byte boundary) and the allocation's size."""
        sf = -(-n // 32) * 32 * 8              # float64 column stride, bytes
        sb = -(-n // 256) * 256                # uint8 column stride, bytes
        nf = len(_FCOLS)
        offs = [i * sf for i in range(nf)] + [nf * sf + j * sb for j in range(len(_BCOLS))]
        return offs, nf * sf + len(_BCOLS) * sb

    @staticmethod
    def _empty(n, coordSys):
        """n rays of uninitialised device columns in one allocation.  The
        column tensors (views) are made on first use; the library reads the
        columns through pointers computed from the layout."""
        n = int(n)
        offs, nbytes = RayVector._layout(n)
        buf = torch.empty(nbytes, dtype=torch.uint8, device=_device.device())
        ret = RayVector._directInit(None, coordSys, buf)
        ret._n = n
        return ret

    @property
    def _cols(self):
        cols = self.__dict__.get("_cols_d")
        if cols is None:                      # views of the single allocation
            n, buf = self._n, self._buf
            offs, _ = RayVector._layout(n)
            cols = {}
            for k, o in zip(_COLS, offs):
                if k in _BCOLS:
                    cols[k] = buf[o: o + n]
                else:
                    cols[k] = buf[o: o + 8 * n].view(torch.float64)
            self.__dict__["_cols_d"] = cols
        return cols

    @_cols.setter
    def _cols(self, cols):
        self.__dict__["_cols_d"] = cols
        if cols is not None:
            self._n = cols["t"].numel()

    def _select(self, minFlux):
        """The rays that are not vignetted and have flux >= minFlux, in order,
        as a new RayVector (the filter of optic.py:1256-1266), by the
        library's two-pass compaction."""
        lib = _lib.load()
        n = len(self)
        nws = int(lib.bt_select_workspace_bytes(n))
        ws = torch.empty(nws + 8, dtype=torch.uint8, device=_device.device())
        count = ws[nws:].view(torch.int64)             # written by bt_select_count
        stream = _device.stream()
        _lib.call("bt_select_count", self._rays, float(minFlux), _device.ptr(count),
                  _device.ptr(ws), nws, stream)
        m = int(count.item())
        out = RayVector._empty(m, self.coordSys)
        if m:
            _lib.call("bt_select_scatter", self._rays, float(minFlux), out._rays,
                      _device.ptr(ws), nws, stream)
        return out

    @staticmethod
    def _count_select(rvs, minFlux):
        """bt_select_count over several RayVectors into one device array
        (``.counts``, no host synchronisation); ``.scatter(i, m)`` then
        copies out the m kept rays of rvs[i] (the filter of
        optic.py:1256-1266)."""
        return _SelectCounts(rvs, minFlux)

    @staticmethod
    def _select_many(rvs, minFlux):
        """``[rv._select(minFlux) for rv in rvs]`` with one host
        synchronisation for all the counts instead of one per RayVector."""
        if not rvs:
            return []
        sel = _SelectCounts(rvs, minFlux)
        return [sel.scatter(i, m) for i, m in enumerate(sel.counts.tolist())]

    @classmethod
    def fromTorch(cls, x, y, z, vx, vy, vz, t, wavelength, flux, vignetted, failed,
                  coordSys=globalCoordSys):
        """Adopt device tensors as columns without copying (they must be
        contiguous float64 / uint8 tensors of equal length on the GPU)."""
        cols = dict(zip(_COLS, (x, y, z, vx, vy, vz, t, wavelength, flux,
                                vignetted, failed)))
        n = cols["x"].numel()
        for k, v in cols.items():
            want = torch.uint8 if k in _BCOLS else torch.float64
            if v.dtype != want or not v.is_cuda or not v.is_contiguous() or v.numel() != n:
                raise ValueError(f"column {k} must be a contiguous {want} CUDA tensor of length {n}")
        return cls._directInit(cols, coordSys)

    # ---- C-ABI view -------------------------------------------------------
    @property
    def _rays(self):
        """The bt_rays descriptor of the columns (built once: the column
        tensors of a RayVector are never replaced, only written in place)."""
        r = self.__dict__.get("_rays_c")
        if r is None:
            r = _lib.BtRays()
            if self.__dict__.get("_cols_d") is None:      # from the layout
                base = self._buf.data_ptr()
                offs, _ = RayVector._layout(self._n)
                for k, o in zip(_COLS, offs):
                    setattr(r, k, base + o)
            else:
                for k in _COLS:
                    setattr(r, k, self._cols[k].data_ptr())
            r.n = self._n
            self._rays_c = r
        return r

    def torch(self, name):
        """The device column ``name`` itself (no copy)."""
        return self._cols[name]

    def __len__(self):
        return self._n

    # ---- host accessors ---------------------------------------------------
    def _host(self, name):
        a = _device.to_numpy(self._cols[name])
        return a.astype(bool) if name in _BCOLS else a

    x = property(lambda self: self._host("x"), doc="x positions [m]")
    y = property(lambda self: self._host("y"), doc="y positions [m]")
    z = property(lambda self: self._host("z"), doc="z positions [m]")
    vx = property(lambda self: self._host("vx"), doc="x velocity [c]")
    vy = property(lambda self: self._host("vy"), doc="y velocity [c]")
    vz = property(lambda self: self._host("vz"), doc="z velocity [c]")
    t = property(lambda self: self._host("t"), doc="optical path length [m]")
    wavelength = property(lambda self: self._host("wavelength"), doc="vacuum wavelength [m]")
    flux = property(lambda self: self._host("flux"), doc="flux")
    vignetted = property(lambda self: self._host("vignetted"), doc="vignetted flags")
    failed = property(lambda self: self._host("failed"), doc="failed flags")

    @property
    def r(self):
        return torch.stack([self._cols[k] for k in ("x", "y", "z")], dim=1).cpu().numpy()

    @property
    def v(self):
        return torch.stack([self._cols[k] for k in ("vx", "vy", "vz")], dim=1).cpu().numpy()

    @property
    def k(self):
        """Wavevectors 2 pi v / (lambda |v|^2) in rad/m (rayVector.py:1178-1190)."""
        v = self.v
        out = 2 * np.pi * v
        out /= self.wavelength[:, None]
        out /= np.sum(v * v, axis=-1)[:, None]
        return out

    kx = property(lambda self: self.k[:, 0])
    ky = property(lambda self: self.k[:, 1])
    kz = property(lambda self: self.k[:, 2])

    @property
    def omega(self):
        return 2 * np.pi / self.wavelength

    # ---- device operations --------------------------------------------------
    def positionAtTime(self, t):
        """Positions at time ``t``, shape (n, 3) (rayVector.cpp:26-54)."""
        n = len(self)
        out = _device.empty_f64(3 * n)
        _lib.call("bt_position_at_time", self._rays, float(t), _device.ptr(out[:n]),
                  _device.ptr(out[n:2 * n]), _device.ptr(out[2 * n:]), _device.stream())
        return _device.to_numpy(out).reshape(3, n).T

    def propagate(self, t):
        """Propagate every ray to time ``t`` in place (rayVector.cpp:56-82)."""
        _lib.call("bt_propagate_in_place", self._rays, float(t), _device.stream())
        return self

    def phase(self, r, t):
        """Plane-wave phases at position ``r`` and time ``t``."""
        out = _device.empty_f64(len(self))
        _lib.call("bt_phase", self._rays, float(r[0]), float(r[1]), float(r[2]),
                  float(t), _device.ptr(out), _device.stream())
        return _device.to_numpy(out)

    def amplitude(self, r, t):
        """Complex amplitudes exp(i phase) at ``r``, ``t``."""
        out = _device.empty_f64(2 * len(self))
        _lib.call("bt_amplitude", self._rays, float(r[0]), float(r[1]), float(r[2]),
                  float(t), _device.ptr(out), _device.stream())
        return _device.to_numpy(out).view(np.complex128)

    def sumAmplitude(self, r, t, ignoreVignetted=True):
        """Flux-weighted sum of amplitudes over unfailed (and, by default,
        unvignetted) rays (rayVector.cpp:169-222)."""
        lib = _lib.load()
        ws = torch.empty(int(lib.bt_sum_amplitude_workspace_bytes(len(self))),
                         dtype=torch.uint8, device=_device.device())
        out = _device.empty_f64(2)
        _lib.call("bt_sum_amplitude", self._rays, float(r[0]), float(r[1]), float(r[2]),
                  float(t), int(bool(ignoreVignetted)), _device.ptr(out), _device.ptr(ws),
                  _device.stream())
        o = _device.to_numpy(out)
        return complex(o[0], o[1])

    def toCoordSys(self, coordSys):
        """Express the rays in ``coordSys`` (in place)."""
        from .trace import applyForwardTransform
        applyForwardTransform(CoordTransform(self.coordSys, coordSys), self)
        return self

    def copy(self):
        out = RayVector._empty(len(self), self.coordSys.copy())
        if self._buf is not None and self._buf.numel() == out._buf.numel():
            out._buf.copy_(self._buf)                  # same layout: one copy
        else:
            for k, v in self._cols.items():
                out._cols[k].copy_(v)
        return out

    def __getitem__(self, idx):
        n = len(self)
        if isinstance(idx, int):
            if idx >= n or idx < -n:
                raise IndexError(f"index {idx} is out of bounds for axis 0 with size {n}")
            idx = slice(idx % n, idx % n + 1)
        if isinstance(idx, slice):
            cols = {k: v[idx].clone() for k, v in self._cols.items()}
        else:
            sel = torch.as_tensor(np.asarray(idx), device=_device.device())
            if sel.dtype == torch.bool:
                sel = torch.nonzero(sel).flatten()
            cols = {k: v[sel].contiguous() for k, v in self._cols.items()}
        return RayVector._directInit(cols, self.coordSys)

    def __eq__(self, rhs):
        if not isinstance(rhs, RayVector) or len(self) != len(rhs):
            return False
        # NaN != NaN, as the reference's elementwise == (dualView.cpp:49-68)
        return all(bool(torch.equal(self._cols[k], rhs._cols[k])) if k in _BCOLS
                   else bool((self._cols[k] == rhs._cols[k]).all())
                   for k in _COLS)

    def __ne__(self, rhs):
        return not (self == rhs)

    def __repr__(self):
        return (f"RayVector({self.x!r}, {self.y!r}, {self.z!r}, {self.vx!r}, {self.vy!r}, "
                f"{self.vz!r}, {self.t!r}, {self.wavelength!r}, {self.flux!r}, "
                f"{self.vignetted!r}, {self.failed!r}, {self.coordSys!r})")

    def __getstate__(self):
        return tuple(self._host(k) for k in _COLS) + (self.coordSys,)

    def __setstate__(self, args):
        self._cols = {k: (_device.u8(a) if k in _BCOLS else _device.f64(a))
                      for k, a in zip(_COLS, args[:-1])}
        self._buf = None
        self._rays_c = None
        self.coordSys = args[-1]

    # ---- generators ----------------------------------------------------------
    @staticmethod
    def _optic_defaults(optic, backDist, medium, stopSurface, coordSys):
        from .optic import Interface
        from .surface import Plane
        if optic is not None:
            if backDist is None:
                backDist = optic.backDist
            if medium is None:
                medium = optic.inMedium
            if stopSurface is None:
                stopSurface = getattr(optic, "stopSurface", None)
            if coordSys is None:
                coordSys = optic.coordSys
        if backDist is None:
            backDist = 40.0
        if stopSurface is None:
            stopSurface = Interface(Plane())
        if medium is None:
            medium = vacuum
        if coordSys is None:
            coordSys = globalCoordSys
        return backDist, medium, stopSurface, coordSys

    @classmethod
    def _generate(cls, gen, npts, stopSurface, coordSys, wavelength, medium,
                  backDist, source, dirCos, flux, out=None):
        """Write ``npts`` new rays on the GPU: stop coordinates from ``gen``
        (a BtRaygen with its pattern fields set), the stop surface's sag,
        the stop -> ``coordSys`` transform (rayVector.py:435-437), then
        finishParallel back to the source plane and the constant columns
        (rayVector.py:816-841), in one bt_generate_rays pass.  ``out``: an
        existing RayVector of ``npts`` rays to overwrite instead of new
        columns."""
        if source is not None:
            raise NotImplementedError("rays from a finite source point are not supported")
        n = medium.getN(wavelength)
        vv = np.array(dirCos, dtype=float, copy=True)
        vv /= n * np.sqrt(np.dot(vv, vv))
        zhat = -n * vv
        xhat = np.cross(np.array([1.0, 0.0, 0.0]), zhat)
        xhat /= np.sqrt(np.dot(xhat, xhat))
        yhat = np.cross(xhat, zhat)
        origin = zhat * backDist
        rot = np.stack([xhat, yhat, zhat]).T
        transform = CoordTransform(stopSurface.coordSys, coordSys)
        gen.stop = stopSurface.surface._dev.ptr()
        gen.stop_dr[:] = [float(v) for v in transform.dr]
        gen.stop_drot[:] = [float(v) for v in np.asarray(transform.drot).ravel()]
        gen.origin[:] = [float(v) for v in origin]
        gen.rot[:] = [float(v) for v in rot.ravel()]
        gen.vv[:] = [float(v) for v in vv]
        gen.wavelength = float(wavelength)
        dev = _device.device()
        fluxd = None
        if isinstance(flux, Real):
            gen.flux0 = float(flux)
        else:
            fluxd = _device.f64(np.broadcast_to(np.asarray(flux, float), (npts,)).ravel())
            gen.flux = fluxd.data_ptr()
        if out is not None:
            if len(out) != npts:
                raise ValueError(f"out holds {len(out)} rays, the generator makes {npts}")
            rv = out
            rv.coordSys = coordSys
        else:
            cols = {k: torch.empty(npts, dtype=torch.float64, device=dev) for k in _FCOLS}
            for k in _BCOLS:
                cols[k] = torch.empty(npts, dtype=torch.uint8, device=dev)
            rv = cls._directInit(cols, coordSys)
        # inputs (tables, points, fluxes) are torch tensors on the current
        # stream: their memory is not reused before this launch has read it
        _lib.call("bt_generate_rays", ctypes.byref(gen), rv._rays, _device.stream())
        del fluxd
        return rv

    @staticmethod
    def _random_pattern(gen, rng, nrandom, lo, span):
        """Point ``gen`` at the PCG64 stream of ``rng`` for two blocks of
        ``nrandom`` uniforms; returns False for other bit generators."""
        words = _pcg.stream_words(rng)
        if words is None:
            return False
        (gen.pcg_state[0], gen.pcg_state[1]), (gen.pcg_inc[0], gen.pcg_inc[1]) = words
        gen.lo[:] = [float(v) for v in lo]
        gen.span[:] = [float(v) for v in span]
        _pcg.consume(rng, 2 * nrandom)
        return True

    @classmethod
    def _points(cls, x, y, *args):
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=float)
        y = np.ascontiguousarray(np.atleast_1d(y), dtype=float)
        gen = _lib.BtRaygen()
        gen.pattern = _lib.GEN_POINTS
        xd, yd = _device.f64(x), _device.f64(y)
        gen.px, gen.py = xd.data_ptr(), yd.data_ptr()
        return cls._generate(gen, x.size, *args)

    @classmethod
    def asGrid(cls, optic=None, backDist=None, medium=None, stopSurface=None,
               coordSys=None, wavelength=None, source=None, dirCos=None,
               theta_x=None, theta_y=None, projection='postel',
               nx=None, ny=None, dx=None, dy=None, lx=None, ly=None,
               flux=1, nrandom=None, rng=None):
        """Rays on a parallelogram grid (or uniformly random within it) on the
        stop surface (rayVector.py:219-444)."""
        if optic is not None and lx is None and (nx is None or dx is None):
            lx = optic.pupilSize
        backDist, medium, stopSurface, coordSys = cls._optic_defaults(
            optic, backDist, medium, stopSurface, coordSys)
        if dirCos is None and source is None:
            dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
        if wavelength is None:
            raise ValueError("Missing wavelength keyword")
        if sum(a is not None for a in [nx, dx, lx]) != 2:
            raise ValueError("Exactly 2 of nx, dx, lx must be specified")
        if nx is not None and ny is None:
            ny = nx
        if dx is not None and dy is None:
            dy = dx
        if lx is not None and ly is None:
            ly = lx if isinstance(lx, Real) else np.dot(np.array([[0, -1], [1, 0]]), lx)
        if nx is not None and dx is not None:
            lx = dx * (nx - 2) if nx % 2 == 0 else dx * (nx - 1)
            ly = dy * (ny - 2) if ny % 2 == 0 else dy * (ny - 1)
        elif lx is not None and dx is not None:
            slop = 0.1
            nx = int((lx / dx + slop) // 2) * 2 + 2
            ny = int((ly / dy + slop) // 2) * 2 + 2
        if isinstance(lx, Real):
            lx = (lx, 0.0)
        if isinstance(ly, Real):
            ly = (0.0, ly)
        args = (stopSurface, coordSys, wavelength, medium, backDist, source, dirCos, flux)
        gen = _lib.BtRaygen()
        gen.lx[:] = [float(v) for v in lx]
        gen.ly[:] = [float(v) for v in ly]
        if nrandom is not None:
            if rng is None:
                rng = np.random.default_rng()
            elif isinstance(rng, int):
                rng = np.random.default_rng(rng)
            nrandom = int(nrandom)
            if cls._random_pattern(gen, rng, nrandom, (-0.5, -0.5), (0.5 - -0.5,) * 2):
                gen.pattern = _lib.GEN_GRID_RANDOM
                return cls._generate(gen, nrandom, *args)
            # another bit generator: its draws come from numpy
            xx = rng.uniform(-0.5, 0.5, size=nrandom)
            yy = rng.uniform(-0.5, 0.5, size=nrandom)
            stack = np.stack([xx, yy])
            return cls._points(np.dot(lx, stack), np.dot(ly, stack), *args)
        # fftshift(fftfreq(n, d)) = (k - n//2) / (n d), k = 0..n-1
        x_d = 1. if nx <= 2 else (nx - (2 if nx % 2 == 0 else 1)) / nx
        y_d = 1. if ny <= 2 else (ny - (2 if ny % 2 == 0 else 1)) / ny
        gen.pattern = _lib.GEN_GRID
        gen.nx, gen.ny = int(nx), int(ny)
        gen.fx = 1.0 / (nx * x_d)
        gen.fy = 1.0 / (ny * y_d)
        return cls._generate(gen, int(nx) * int(ny), *args)

    @classmethod
    def asPolar(cls, optic=None, backDist=None, medium=None, stopSurface=None,
                coordSys=None, wavelength=None, outer=None, inner=None,
                source=None, dirCos=None, theta_x=None, theta_y=None,
                projection='postel', nrad=None, naz=None, kfold=6, flux=1,
                nrandom=None, rng=None):
        """Rays on a hexapolar grid (or uniformly random) over an annulus of
        the stop surface (rayVector.py:459-643)."""
        if optic is not None:
            if outer is None:
                outer = optic.pupilSize / 2
            if inner is None:
                inner = (optic.pupilSize * optic.pupilObscuration / 2
                         if hasattr(optic, 'pupilObscuration') else 0.0)
        elif inner is None:
            inner = 0.0
        backDist, medium, stopSurface, coordSys = cls._optic_defaults(
            optic, backDist, medium, stopSurface, coordSys)
        if dirCos is None and source is None:
            dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
        if wavelength is None:
            raise ValueError("Missing wavelength keyword")
        args = (stopSurface, coordSys, wavelength, medium, backDist, source, dirCos, flux)
        gen = _lib.BtRaygen()
        if nrandom is None:
            npts, _tables = cls._polar_tables(gen, outer, inner, nrad, naz, kfold)
            return cls._generate(gen, npts, *args)
        if rng is None:
            rng = np.random.default_rng()
        elif isinstance(rng, int):
            rng = np.random.default_rng(rng)
        nrandom = int(nrandom)
        lo0, hi0 = inner**2, outer**2
        if cls._random_pattern(gen, rng, nrandom, (lo0, 0.0), (hi0 - lo0, 2 * np.pi - 0.0)):
            gen.pattern = _lib.GEN_POLAR_RANDOM
            return cls._generate(gen, nrandom, *args)
        rr = np.sqrt(rng.uniform(inner**2, outer**2, size=nrandom))
        th = rng.uniform(0, 2 * np.pi, size=nrandom)
        return cls._points(rr * np.cos(th), rr * np.sin(th), *args)

    @staticmethod
    def _polar_tables(gen, outer, inner, nrad, naz, kfold=6):
        """Hexapolar rings (utils.py:512-531) as the device ring table of a
        POLAR generator: ray j of ring k sits at th = j * (2 pi / nphi_k),
        numpy's linspace(0, 2 pi, nphi_k, endpoint=False).  Returns the ray
        count and the device tables (keep them alive until the launch)."""
        rhos, nphis, dth = _hexapolar_rings(outer, inner, nrad, naz, kfold)
        off = np.zeros(len(nphis) + 1, dtype=np.int64)
        off[1:] = np.cumsum(nphis)
        off_d = torch.from_numpy(off).to(_device.device())
        rho_d, dth_d = _device.f64(rhos), _device.f64(dth)
        gen.pattern = _lib.GEN_POLAR
        gen.nring = len(nphis)
        gen.ring_off, gen.ring_rho, gen.ring_dth = (
            off_d.data_ptr(), rho_d.data_ptr(), dth_d.data_ptr())
        return int(off[-1]), (off_d, rho_d, dth_d)

    @classmethod
    def asSpokes(cls, optic=None, backDist=None, medium=None, stopSurface=None,
                 coordSys=None, wavelength=None, outer=None, inner=0.0,
                 source=None, dirCos=None, theta_x=None, theta_y=None,
                 projection='postel', spokes=None, rings=None, spacing='uniform',
                 flux=1):
        """Rays on rings x spokes, optionally at Gauss-quadrature radii with
        quadrature weights as fluxes (rayVector.py:645-814)."""
        from numbers import Integral
        if optic is not None and outer is None:
            outer = optic.pupilSize / 2
        backDist, medium, stopSurface, coordSys = cls._optic_defaults(
            optic, backDist, medium, stopSurface, coordSys)
        if dirCos is None and source is None:
            dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
        if wavelength is None:
            raise ValueError("Missing wavelength keyword")
        if isinstance(rings, Integral):
            if spacing == 'uniform':
                rings = np.linspace(inner, outer, rings)
            elif spacing == 'GQ':
                if spokes is None:
                    spokes = 2 * rings + 1
                Li, w = np.polynomial.legendre.leggauss(rings)
                eps = inner / outer
                area = np.pi * (1 - eps**2)
                rings = np.sqrt(eps**2 + (1 + Li) * (1 - eps**2) / 2) * outer
                flux = w * area / (2 * spokes)
        if isinstance(spokes, Integral):
            spokes = np.linspace(0, 2 * np.pi, spokes, endpoint=False)
        rings, spokes = np.meshgrid(rings, spokes)
        flux = np.broadcast_to(flux, rings.shape).ravel()
        rings = rings.ravel()
        spokes = spokes.ravel()
        return cls._points(rings * np.cos(spokes), rings * np.sin(spokes), stopSurface,
                           coordSys, wavelength, medium, backDist, source, dirCos, flux)

    @classmethod
    def fromStop(cls, x, y, optic=None, backDist=None, medium=None,
                 stopSurface=None, coordSys=None, wavelength=None, source=None,
                 dirCos=None, theta_x=None, theta_y=None, projection='postel',
                 flux=1):
        """Rays that would cross the stop surface at (x, y)
        (rayVector.py:860-982)."""
        backDist, medium, stopSurface, coordSys = cls._optic_defaults(
            optic, backDist, medium, stopSurface, coordSys)
        if dirCos is None and source is None:
            dirCos = fieldToDirCos(theta_x, theta_y, projection=projection)
        if wavelength is None:
            raise ValueError("Missing wavelength keyword")
        return cls._points(x, y, stopSurface, coordSys, wavelength, medium, backDist,
                           source, dirCos, flux)

    @classmethod
    def asFan(cls, nx=None, ny=None, **kwargs):
        rvs = []
        if nx > 1:
            rvs.append(RayVector.asGrid(nx=nx, ny=1, **kwargs))
        if ny > 1:
            rvs.append(RayVector.asGrid(nx=1, ny=ny, **kwargs))
        return concatenateRayVectors(rvs)


def concatenateRayVectors(rvs):
    """Rays of several RayVectors in one (rayVector.py:1351-1365)."""
    cols = {k: torch.cat([rv._cols[k] for rv in rvs]) for k in _COLS}
    return RayVector._directInit(cols, rvs[0].coordSys)


class _SelectCounts:
    """bt_select_count launched over several ray sets, their counts in one
    device int64 array and their per-block offsets kept for the copies."""

    def __init__(self, rvs, minFlux):
        lib = _lib.load()
        self.rvs = rvs
        self.minFlux = float(minFlux)
        self.sizes = [int(lib.bt_select_workspace_bytes(len(rv))) for rv in rvs]
        self.offs, tot = [], 0
        for b in self.sizes:
            self.offs.append(tot)
            tot += -(-b // 256) * 256
        self.ws = torch.empty(tot + 8 * len(rvs), dtype=torch.uint8, device=_device.device())
        self.counts = self.ws[tot:].view(torch.int64)   # written by bt_select_count
        stream = _device.stream()
        for i, rv in enumerate(rvs):
            _lib.call("bt_select_count", rv._rays, self.minFlux, _device.ptr(self.counts[i:]),
                      _device.ptr(self.ws[self.offs[i]:]), self.sizes[i], stream)

    def scatter(self, i, m):
        rv = self.rvs[i]
        out = RayVector._empty(m, rv.coordSys)
        if m:
            _lib.call("bt_select_scatter", rv._rays, self.minFlux, out._rays,
                      _device.ptr(self.ws[self.offs[i]:]), self.sizes[i], _device.stream())
        return out
