"""Device-memory plumbing: torch tensors on the GPU hold ray columns, tables
and descriptor records; the HIP library receives their raw pointers."""
import ctypes

import numpy as np
import torch

from . import _lib


_ready = False


def device():
    global _ready
    if not _ready:
        if not torch.cuda.is_available():
            raise _lib.BatoidAMDError(
                "batoid_amd needs a ROCm GPU (torch.cuda.is_available() is False)"
            )
        _lib.load()
        _ready = True
    return torch.device("cuda", torch.cuda.current_device())


# torch's raw current-stream query (what its own generated code calls) when it
# has one; the public Stream object otherwise
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream():
    """Handle of torch's current HIP stream: library calls are ordered with
    torch's own work on it."""
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def f64(a):
    """A contiguous float64 device copy of array-like ``a``."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=torch.float64).contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if not arr.flags.writeable:
        arr = arr.copy()
    return torch.from_numpy(arr).to(device())


def i32(a):
    """A contiguous int32 device copy of numpy array ``a``."""
    arr = np.ascontiguousarray(a, dtype=np.int32)
    if not arr.flags.writeable:
        arr = arr.copy()
    return torch.from_numpy(arr).to(device())


def u8(a):
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=torch.uint8).contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=bool).astype(np.uint8))
    return torch.from_numpy(arr).to(device())


def upload_bytes(b):
    """Upload raw descriptor bytes; returns a uint8 device tensor."""
    if len(b) == 0:
        return torch.empty(0, dtype=torch.uint8, device=device())
    host = torch.frombuffer(bytearray(b), dtype=torch.uint8)
    return host.to(device())


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def empty_f64(n):
    return torch.empty(int(n), dtype=torch.float64, device=device())


def to_numpy(t):
    return t.detach().cpu().numpy()


class Materialized:
    """Descriptor records uploaded to device memory, with their tables.

    ``records`` is a list of ctypes structs whose pointer fields a0..a3 are
    filled here from the matching numpy arrays in ``arrays``.  The uploaded
    buffer and every table stay referenced by this object."""

    def __init__(self, records, arrays):
        self.tables = []
        recs = []
        for rec, arrs in zip(records, arrays):
            rec = type(rec).from_buffer_copy(rec)
            for k, a in enumerate(arrs):
                if a is None:
                    continue
                t = i32(a) if getattr(a, "dtype", None) == np.int32 else f64(a)
                self.tables.append(t)
                setattr(rec, f"a{k}", t.data_ptr())
            recs.append(rec)
        self.count = len(recs)
        self.buffer = upload_bytes(_lib.struct_array_bytes(recs))
        self.record_size = ctypes.sizeof(type(records[0])) if records else 0

    def ptr(self, index=0):
        return ctypes.c_void_p(self.buffer.data_ptr() + index * self.record_size)
