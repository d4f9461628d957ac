"""Tile-variant choice for the segmented reductions, and the calls of the
library's tuned/fused entry points (``awkward_hip_reduce_*``).

The variants (threads, lists per thread, LDS staging bytes) are the ones
compiled into libawkward_hip.so (``csrc/reduce.h``, ``launch_reduce``).  The
fixed heuristic below models each variant's expected staged span (lists x
mean list length x element size) and occupancy (workgroups per CU from LDS
and wave limits) and picks the one with the most bytes in flight per CU
among those whose stage should hold the span; the mean list length is
elements / lists, which the host knows from the array shapes without
synchronising with the device.  ``awkward_amd.tuning`` replaces the
heuristic by measured choices (the autotuner table, or the learned
dispatcher).
"""

from __future__ import annotations

import ctypes

import numpy as np

VARIANTS = [
    (256, 1, 32768),
    (256, 4, 8192),
    (256, 2, 16384),
    (128, 1, 16384),
    (256, 1, 65536),
    (256, 8, 8192),
    (64, 1, 6144),
    (128, 2, 16384),
    (64, 4, 8192),
    (128, 1, 8192),
]

LDS_PER_CU = 160 * 1024
WAVES_PER_CU = 32

#: optional override: callable(op, n_lists, n_elements, itemsize, offsets) -> variant
#: (``awkward_amd.learned.enable()`` installs the learned dispatcher here)
chooser = None


def variant_model(v: int, n_lists: int, n_elements: int, itemsize: int):
    """(expected fill of the LDS stage, expected bytes in flight per CU)."""
    block, lpt, cap = VARIANTS[v]
    mean = n_elements / max(n_lists, 1)
    lists = block * lpt
    staged = lists * mean * itemsize
    lds = cap + 8 * (lists + 1)
    per_cu = min(LDS_PER_CU // lds, WAVES_PER_CU // max(block // 64, 1))
    per_cu = max(per_cu, 1)
    inflight = per_cu * (min(staged, cap) + 8 * (lists + 1))
    return staged / cap, inflight


def heuristic_variant(n_lists: int, n_elements: int, itemsize: int) -> int:
    """The variant with the most bytes in flight per CU among those whose
    stage is expected to hold the workgroup's span (fill <= 85%)."""
    best, best_key = None, None
    for v in range(len(VARIANTS)):
        fill, inflight = variant_model(v, n_lists, n_elements, itemsize)
        if fill > 0.85:
            continue
        key = (inflight, -v)
        if best_key is None or key > best_key:
            best, best_key = v, key
    return 4 if best is None else best  # long lists: the largest stage, heavy paths


#: the long-list queue (``csrc/reduce.h``, ``kLongLists``): chunks of very long
#: lists spread over the chip instead of one workgroup per list
LONG_LISTS = 10
#: mean list length from which the long-list queue is used
LONG_MEAN = 2048


def choose_variant(op: str, n_lists: int, n_elements: int, itemsize: int, offsets=None) -> int:
    if n_lists > 0 and n_elements >= LONG_MEAN * n_lists:
        return LONG_LISTS
    if chooser is not None:
        return int(chooser(op, n_lists, n_elements, itemsize, offsets))
    return heuristic_variant(n_lists, n_elements, itemsize)


_TYPE = {
    "int8": ctypes.c_int8, "uint8": ctypes.c_uint8, "int16": ctypes.c_int16, "uint16": ctypes.c_uint16,
    "int32": ctypes.c_int32, "uint32": ctypes.c_uint32, "int64": ctypes.c_int64, "uint64": ctypes.c_uint64,
    "float32": ctypes.c_float, "float64": ctypes.c_double,
}
_fns = {}


def _fn(name, argtypes):
    f = _fns.get(name)
    if f is None:
        from awkward_amd._kernels import KernelError
        from awkward_amd._lib import lib

        f = getattr(lib(), name)
        f.argtypes = argtypes
        f.restype = KernelError
        _fns[name] = f
    return f


P = ctypes.c_void_p
I64 = ctypes.c_int64
INT = ctypes.c_int


def reduce_sum(out, data, offsets, n, variant):
    from awkward_amd._kernels import device_address, set_stream

    on = np.dtype(_np(out.dtype)).name
    inn = np.dtype(_np(data.dtype)).name
    f = _fn(f"awkward_hip_reduce_sum_{on}_{inn}", [P, P, P, I64, INT])
    set_stream()
    return f(device_address(out), device_address(data), device_address(offsets), n, variant)


def reduce_prod(out, data, offsets, n, variant):
    from awkward_amd._kernels import device_address, set_stream

    on = np.dtype(_np(out.dtype)).name
    inn = np.dtype(_np(data.dtype)).name
    f = _fn(f"awkward_hip_reduce_prod_{on}_{inn}", [P, P, P, I64, INT])
    set_stream()
    return f(device_address(out), device_address(data), device_address(offsets), n, variant)


def reduce_anyall(kind, out, mask, data, offsets, n, variant):
    """kind "any" (awkward_reduce_sum_bool) or "all" (awkward_reduce_prod_bool)."""
    from awkward_amd._kernels import device_address, set_stream

    tn = np.dtype(_np(data.dtype)).name
    f = _fn(f"awkward_hip_reduce_{kind}_masked_{tn}", [P, P, P, P, I64, INT])
    set_stream()
    return f(device_address(out), 0 if mask is None else device_address(mask), device_address(data),
             device_address(offsets), n, variant)


def reduce_extremum(kind, out, mask, data, offsets, n, identity, variant):
    from awkward_amd._kernels import device_address, set_stream

    tn = np.dtype(_np(data.dtype)).name
    f = _fn(f"awkward_hip_reduce_{kind}_masked_{tn}", [P, P, P, P, I64, _TYPE[tn], INT])
    set_stream()
    return f(device_address(out), 0 if mask is None else device_address(mask), device_address(data),
             device_address(offsets), n, identity, variant)


def reduce_arg(kind, out, mask, data, offsets, n, variant, shifts=None, starts=None):
    """Local argmin/argmax; with ``shifts`` (below an option type) the index
    is made local to ``starts`` (the lists' starts before the nulls were
    compacted away) instead of to ``offsets``."""
    from awkward_amd._kernels import device_address, set_stream

    tn = np.dtype(_np(data.dtype)).name
    set_stream()
    m = 0 if mask is None else device_address(mask)
    if shifts is not None:
        f = _fn(f"awkward_hip_reduce_{kind}_shifts_{tn}", [P, P, P, P, P, P, I64, INT])
        return f(device_address(out), m, device_address(data), device_address(offsets), device_address(starts),
                 device_address(shifts), n, variant)
    f = _fn(f"awkward_hip_reduce_{kind}_local_{tn}", [P, P, P, P, I64, INT])
    return f(device_address(out), m, device_address(data), device_address(offsets), n, variant)


def _np(dtype):
    from awkward_amd.nplike import to_numpy_dtype

    return to_numpy_dtype(dtype)
