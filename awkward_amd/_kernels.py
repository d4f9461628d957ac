"""Kernel objects of the hip backend.

Mirrors the reference's kernel-call contract (``src/awkward/_kernels.py``):
a kernel is looked up by ``(name, *dtypes)`` and called with the
specification's arguments in order -- arrays for pointer arguments, Python
scalars for the rest -- and returns an error struct with ``str``,
``filename``, ``id`` and ``attempt`` (``_kernels.py:28-42``).  Arrays here are
torch tensors resident on the current ROCm device; the call passes their
device addresses to ``libawkward_hip.so`` and enqueues on torch's current
stream.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from awkward_amd._kernel_signatures import SPECIALIZATIONS
from awkward_amd._lib import lib


class KernelError(ctypes.Structure):
    _fields_ = [
        ("str", ctypes.c_char_p),
        ("filename", ctypes.c_char_p),
        ("id", ctypes.c_int64),
        ("attempt", ctypes.c_int64),
    ]


_SCALAR = {
    "bool": ctypes.c_bool,
    "int8_t": ctypes.c_int8,
    "uint8_t": ctypes.c_uint8,
    "int16_t": ctypes.c_int16,
    "uint16_t": ctypes.c_uint16,
    "int32_t": ctypes.c_int32,
    "uint32_t": ctypes.c_uint32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "float": ctypes.c_float,
    "double": ctypes.c_double,
}


def _argtype(spec_type: str):
    return ctypes.c_void_p if "List[" in spec_type else _SCALAR[spec_type]


def set_stream() -> None:
    """Point the library at torch's current HIP stream."""
    lib().awkward_hip_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


class PointerArray:
    """A device array of device pointers (the combinations kernel's tocarry)."""

    def __init__(self, tensors):
        self.tensors = list(tensors)
        addrs = np.array([t.data_ptr() for t in self.tensors], dtype=np.uint64)
        device = self.tensors[0].device if self.tensors else torch.device("cuda")
        self.table = torch.from_numpy(addrs.view(np.int64)).to(device)

    def data_ptr(self):
        return self.table.data_ptr()


def device_address(x) -> int:
    if isinstance(x, PointerArray):
        return x.data_ptr()
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise TypeError("hip kernels take device tensors; got a CPU tensor")
        if not x.is_contiguous():
            raise TypeError("hip kernels expect contiguous arrays")
        return x.data_ptr()
    raise TypeError(f"hip kernels take device tensors, not {type(x).__name__}")


class HipKernel:
    def __init__(self, name: str):
        family, args, key = SPECIALIZATIONS[name]
        self.name = name
        self.family = family
        self.key = key
        self.dir = [d for _, _, d in args]
        self._pointer = ["List[" in t for _, t, _ in args]
        self._impl = getattr(lib(), name)
        self._impl.argtypes = [_argtype(t) for _, t, _ in args]
        self._impl.restype = KernelError

    def __repr__(self):
        return f"<HipKernel {self.name}>"

    def __call__(self, *args) -> KernelError:
        assert len(args) == len(self._pointer), (self.name, len(args))
        call = [device_address(a) if p else a for a, p in zip(args, self._pointer)]
        set_stream()
        return self._impl(*call)
