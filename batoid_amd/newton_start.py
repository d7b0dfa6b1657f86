"""Learned Newton starts for the fused trace.

No counterpart in the reference.  The reference intersects a Newton surface
(Asphere, Bicubic, Zernike/PolynomialSurface, Sum) by starting from an
analytic root -- the Asphere's conic (asphere.cpp:75-85) or a Sum's first
child (sum.cpp:47-56), the ray origin for a polynomial or Bicubic surface
(surface.cpp:22-49) -- and iterating Newton a fixed number of times.  Here a
small network per Newton surface

    6 -> 64 -> 64 -> 1, ReLU

predicts the residual ``dt* - dt_conic`` between the exact intersection time
and the surface's base root (the conic's root; 0, i.e. the ray origin, for
surfaces without a conic base) from the ray's point at that root and its
direction, in the surface's local frame.  Inside the fused trace each
network is an fp16 MFMA product evaluated by whole waves at its step.
Training data come from the tracer (``bt_newton_samples``); training runs on
the GPU in FP32 with the library's own forward/backward and Adam kernels.

The start changes only where Newton begins: every ray is still converged on
the full surface (to a sag gap of 1e-14 m) and any ray that does not
converge from the learned start is re-intersected on the reference path.
Networks are specific to the surface and the ray bundles they were trained
on.

Usage::

    nets = train_newton_starts(optic, [rv1, rv2, ...])
    optic.trace(rv, start=nets)        # or start="conic", start="reference"
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib, _device
from .coordTransform import CoordTransform

N_IN = 6
N_HID = 64


def param_count():
    return _lib.load().bt_mlp_param_count()


def _fptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _is_newton(surface):
    """Surfaces the reference intersects by Newton iteration."""
    recs, _ = surface._records()
    return recs[0].kind in (_lib.ASPHERE, _lib.BICUBIC, _lib.POLYNOMIAL, _lib.SUM)


class NewtonStartNet:
    """The network of one surface: FP32 parameters on the GPU, input and
    output normalisation, and the packed fp16 tables the trace reads."""

    def __init__(self, seed=0):
        rng = np.random.default_rng(seed)
        p = np.zeros(param_count(), dtype=np.float32)
        o = 0
        p[o:o + N_HID * N_IN] = rng.normal(0, math.sqrt(2.0 / N_IN), N_HID * N_IN)
        o += N_HID * N_IN + N_HID
        p[o:o + N_HID * N_HID] = rng.normal(0, math.sqrt(2.0 / N_HID), N_HID * N_HID)
        o += N_HID * N_HID + N_HID
        p[o:o + N_HID] = rng.normal(0, math.sqrt(1.0 / N_HID), N_HID)
        self.params = torch.from_numpy(p).to(_device.device())
        self.in_norm = np.concatenate([np.zeros(N_IN), np.ones(N_IN)]).astype(np.float32)
        self.out_mean = 0.0
        self.out_std = 1.0
        self.history = []
        self.samples = 0
        self._packed = None

    # -- normalisation ------------------------------------------------------
    def _norm_dev(self):
        return torch.from_numpy(self.in_norm.astype(np.float32)).to(_device.device())

    def set_normalisation(self, features, delta):
        f = features.double()
        mean = f.mean(dim=0).cpu().numpy()
        std = f.std(dim=0).cpu().numpy()
        std = np.where(std > 0, std, 1.0)
        self.in_norm = np.concatenate([mean, 1.0 / std]).astype(np.float32)
        d = delta.double()
        self.out_mean = float(d.mean())
        s = float(d.std())
        self.out_std = s if s > 0 else 1.0
        self._packed = None

    # -- training -----------------------------------------------------------
    def loss_grad(self, features, targets, index=None):
        """(loss, grad) of the mean squared error in normalised units over
        ``features[index]`` (all rows when index is None)."""
        n = int(index.numel()) if index is not None else int(features.shape[0])
        ws_bytes = _lib.load().bt_mlp_grad_workspace_bytes(n)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=features.device)
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=features.device)
        norm = self._norm_dev()
        _lib.call("bt_mlp_loss_grad", _fptr(self.params), _fptr(features), _fptr(targets),
                  _device.ptr(index), n, _fptr(norm), _fptr(grad), _fptr(loss), _fptr(ws),
                  ws_bytes, _device.stream())
        return loss, grad

    def fit(self, features, delta, steps=4000, batch=8192, lr=3e-3, lr_final=3e-5,
            seed=0, log_every=250):
        """Train on ``features`` [n, 6] (float32, device) against ``delta``
        [n] (dt* - dt_conic, float64 metres).  Adam with a cosine learning
        rate from ``lr`` to ``lr_final``; minibatches drawn with replacement.
        Returns the (step, minibatch loss) history, logged at step 1 and every
        ``log_every`` steps (the loss is the one before that step's update)."""
        dev = _device.device()
        features = features.to(dev, torch.float32).contiguous()
        delta = delta.to(dev, torch.float64)
        n = features.shape[0]
        if n == 0:
            raise ValueError("no training samples")
        self.samples = n
        self.set_normalisation(features, delta)
        targets = ((delta - self.out_mean) / self.out_std).to(torch.float32).contiguous()
        batch = min(batch, n)
        lib = _lib.load()
        ws_bytes = lib.bt_mlp_grad_workspace_bytes(batch)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        grad = torch.empty_like(self.params)
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=dev)
        norm = self._norm_dev()
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        st = _device.stream()
        self.history = []
        for step in range(1, steps + 1):
            idx = torch.randint(0, n, (batch,), generator=gen, device=dev, dtype=torch.int32)
            _lib.call("bt_mlp_loss_grad", _fptr(self.params), _fptr(features), _fptr(targets),
                      _fptr(idx), batch, _fptr(norm), _fptr(grad), _fptr(loss), _fptr(ws),
                      ws_bytes, st)
            frac = (step - 1) / max(steps - 1, 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * frac))
            _lib.call("bt_mlp_adam", _fptr(self.params), _fptr(grad), _fptr(m), _fptr(v),
                      self.params.numel(), rate, 0.9, 0.999, 1e-8, step, st)
            if log_every and (step == 1 or step % log_every == 0 or step == steps):
                self.history.append((step, float(loss.item())))
        self._packed = None
        return self.history

    # -- inference ----------------------------------------------------------
    def predict(self, features):
        """FP32 prediction of dt* - dt_conic [m] (float64 tensor)."""
        features = features.to(_device.device(), torch.float32).contiguous()
        n = features.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=features.device)
        _lib.call("bt_mlp_forward", _fptr(self.params), _fptr(self._norm_dev()),
                  _fptr(features), n, _fptr(out), _device.stream())
        return out.double() * self.out_std + self.out_mean

    @property
    def packed(self):
        """The fp16 MFMA tables read by the trace (uint8 device tensor)."""
        if self._packed is None:
            lib = _lib.load()
            buf = torch.empty(lib.bt_mlp_packed_bytes(), dtype=torch.uint8,
                              device=_device.device())
            norm = np.concatenate([self.in_norm, [self.out_mean, self.out_std]]).astype(np.float32)
            normd = torch.from_numpy(norm).to(buf.device)
            _lib.call("bt_mlp_pack", _fptr(self.params), _fptr(normd), _fptr(buf),
                      _device.stream())
            self._packed = buf
        return self._packed

    def predict_packed(self, features):
        """fp16-MFMA prediction, as evaluated inside the trace [m] (float32)."""
        features = features.to(_device.device(), torch.float32).contiguous()
        n = features.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=features.device)
        _lib.call("bt_mlp_forward_packed", _fptr(self.packed), _fptr(features), n,
                  _fptr(out), _device.stream())
        return out

    # -- persistence --------------------------------------------------------
    def state(self):
        return {"params": self.params.cpu().numpy(), "in_norm": self.in_norm.copy(),
                "out_mean": self.out_mean, "out_std": self.out_std}

    @classmethod
    def from_state(cls, state):
        net = cls.__new__(cls)
        net.params = torch.from_numpy(np.asarray(state["params"], dtype=np.float32)).to(
            _device.device())
        net.in_norm = np.asarray(state["in_norm"], dtype=np.float32)
        net.out_mean = float(state["out_mean"])
        net.out_std = float(state["out_std"])
        net.history = []
        net.samples = 0
        net._packed = None
        return net


def newton_samples(surface, coordSys, rv):
    """Training samples for rays ``rv`` (in their own frame) meeting
    ``surface`` placed at ``coordSys``: features [n, 6] float32, dt_conic and
    dt_exact [n] float64, status [n] uint8 (bit 0: conic root found, bit 1:
    reference intersection converged).  ``rv`` is not modified."""
    from .global_vars import _batoid_niter
    ct = CoordTransform(rv.coordSys, coordSys)
    n = len(rv)
    dev = _device.device()
    feat = torch.empty((n, N_IN), dtype=torch.float32, device=dev)
    dtc = torch.empty(n, dtype=torch.float64, device=dev)
    dte = torch.empty(n, dtype=torch.float64, device=dev)
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    _lib.call("bt_newton_samples", surface._dev.ptr(), _lib.dvec(ct.dr),
              _lib.dvec(np.asarray(ct.drot).ravel()), rv._rays, _batoid_niter, _fptr(feat),
              _fptr(dtc), _fptr(dte), _fptr(status), _device.stream())
    return feat, dtc, dte, status


def collect_samples(optic, rvs, reverse=False):
    """Trace each RayVector of ``rvs`` through ``optic`` one Interface at a
    time and record, at every Newton surface, the samples of the rays that
    reach it.  Returns {interface name: (features, delta)}."""
    from .plan import collect_steps
    steps = collect_steps(optic, reverse)
    acc = {}
    for rv in rvs:
        rv = rv.copy()
        for iface, _, _ in steps:
            if _is_newton(iface.surface):
                feat, dtc, dte, status = newton_samples(iface.surface, iface.coordSys, rv)
                ok = status == 3
                acc.setdefault(iface.name, []).append((feat[ok], (dte - dtc)[ok]))
            iface.trace(rv, reverse=reverse)
    return {name: (torch.cat([f for f, _ in parts]), torch.cat([d for _, d in parts]))
            for name, parts in acc.items()}


def train_newton_starts(optic, rvs, reverse=False, seed=0, **fit_kwargs):
    """A trained NewtonStartNet for every Newton surface of ``optic`` that
    the rays ``rvs`` reach: {interface name: net}."""
    nets = {}
    for k, (name, (feat, delta)) in enumerate(sorted(collect_samples(optic, rvs, reverse).items())):
        if feat.shape[0] == 0:
            continue
        net = NewtonStartNet(seed=seed + k)
        net.fit(feat, delta, seed=seed + k, **fit_kwargs)
        nets[name] = net
    return nets
