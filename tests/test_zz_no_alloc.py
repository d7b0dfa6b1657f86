"""bt_huygens_gemm_noalloc -- the allocation-free form of bt_huygens_gemm --
called through the C ABI: the same images, bit for bit, as bt_huygens_gemm
with a caller workspace, on two calls in a row (the second with fewer rays
and a smaller image, so stale scratch from the first call would show), in
both MFMA precisions; requests beyond the fixed scratch are refused with
BT_E_SIZE before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

import batoid_amd as bt
from batoid_amd import _lib, _device, workloads
from batoid_amd.analysis import _chief_index
from batoid_amd.utils import fieldToDirCos

pytestmark = pytest.mark.gpu

PREC = {"fp32": _lib.PSF_FP32, "bf16x3": _lib.PSF_BF16X3}


def _traced(nx, field=0):
    optic = workloads.tma()
    thx, thy = workloads.tma_fields()
    rays = bt.RayVector.asGrid(optic=optic, wavelength=620e-9,
                               dirCos=fieldToDirCos(thx[field], thy[field]), nx=nx)
    optic.trace(rays)
    return rays, rays.r[_chief_index(nx)], float(rays.t[0])


def _with_workspace(rays, point, prim, nx, ny, time, prec):
    lib = _lib.load()
    nbytes = int(lib.bt_huygens_workspace_bytes(len(rays), nx, ny))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=_device.device())
    out = _device.empty_f64(2 * nx * ny)
    _lib.check(lib.bt_huygens_gemm(rays._rays, _lib.dvec(point[:2]), _lib.dvec(prim[0]),
                                   _lib.dvec(prim[1]), nx, ny, time, 1, PREC[prec],
                                   _device.ptr(out), _device.ptr(ws), nbytes,
                                   _device.stream()), "bt_huygens_gemm")
    return _device.to_numpy(out)


def _noalloc(rays, point, prim, nx, ny, time, prec):
    lib = _lib.load()
    out = _device.empty_f64(2 * nx * ny)
    _lib.check(lib.bt_huygens_gemm_noalloc(rays._rays, _lib.dvec(point[:2]),
                                           _lib.dvec(prim[0]), _lib.dvec(prim[1]), nx, ny,
                                           time, 1, PREC[prec], _device.ptr(out),
                                           _device.stream()), "bt_huygens_gemm_noalloc")
    return _device.to_numpy(out)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_noalloc_equals_workspace_call_twice(prec):
    print(f"case noalloc {prec}: 512^2 grid -> 256^2, then 96^2 grid -> 72 x 40",
          flush=True)
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    big, p_big, t_big = _traced(512)
    small, p_small, t_small = _traced(96, field=5)
    a1 = _noalloc(big, p_big, prim, 256, 256, t_big, prec)
    b1 = _with_workspace(big, p_big, prim, 256, 256, t_big, prec)
    np.testing.assert_array_equal(a1.view(np.uint64), b1.view(np.uint64))
    assert np.abs(a1).max() > 0
    sheared = np.array([[0.6e-6, 0.1e-6], [-0.05e-6, 0.55e-6]])
    a2 = _noalloc(small, p_small, sheared, 72, 40, t_small, prec)
    b2 = _with_workspace(small, p_small, sheared, 72, 40, t_small, prec)
    np.testing.assert_array_equal(a2.view(np.uint64), b2.view(np.uint64))
    assert np.abs(a2).max() > 0


def test_noalloc_refuses_requests_beyond_the_scratch():
    lib = _lib.load()
    cap = int(lib.bt_huygens_noalloc_scratch_bytes())
    assert cap == 128 << 20
    rays = bt.RayVector(np.zeros(8), 0.0, 0.0, 0.0, 0.0, 1.0)
    r = rays._rays
    big = _lib.BtRays.from_buffer_copy(r)
    big.n = 1 << 22                      # 4 Mi rays: 400 MB of per-ray parameters
    out = _device.empty_f64(2 * 16 * 16)
    rc = lib.bt_huygens_gemm_noalloc(big, _lib.dvec([0, 0]), _lib.dvec([1e-6, 0]),
                                     _lib.dvec([0, 1e-6]), 16, 16, 0.0, 1, _lib.PSF_FP32,
                                     _device.ptr(out), _device.stream())
    assert rc == -2                      # BT_E_SIZE
    assert b"scratch" in lib.bt_last_error()
    assert int(lib.bt_huygens_workspace_bytes(1 << 22, 16, 16)) > cap
