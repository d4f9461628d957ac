"""Batch runner: many independent jagged arrays, one persistent launch per
dtype and op, capturable as a single hipGraph.

``BatchRunner(arrays)`` takes (offsets, content) pairs of torch tensors on
the device (int64 zero-based offsets; int32/int64/float32/float64 content),
groups them by dtype and, per group, builds a device table of the arrays and
their tile numbering.  ``run()`` enqueues, per group, a counter reset, the
fused sum/min/argmin pass (``awkward_hip_batch_reduce_*``) and the argsort
(``awkward_hip_batch_argsort_*``); no host synchronisation, so
``awkward_amd.fused.Graph(runner.run)`` captures the whole batch.

The per-array results match the registry path on the same array:
``ak.sum``, ``ak.min`` / ``ak.argmin`` (mask_identity=True: value + empty
mask) and ``ak.argsort`` (stable, ascending).

``variant`` selects the reduce tile per dtype (see ``REDUCE_VARIANTS``); the
autotuner (``tune()``) times them on the batch itself and keeps the fastest.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from awkward_amd._kernels import KernelError, set_stream
from awkward_amd._lib import lib

#: (threads, lists per thread, LDS bytes) of awkward_hip_batch_reduce_* variants
REDUCE_VARIANTS = [(128, 1, 16384), (64, 1, 8192), (256, 4, 8192), (256, 8, 8192), (256, 1, 32768)]

_SUM = {np.dtype(np.int32): torch.int64, np.dtype(np.int64): torch.int64,
        np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}
_NAME = {np.dtype(np.int32): "int32", np.dtype(np.int64): "int64", np.dtype(np.float32): "float32",
         np.dtype(np.float64): "float64"}
_CT = {"int32": ctypes.c_int32, "int64": ctypes.c_int64, "float32": ctypes.c_float, "float64": ctypes.c_double}
_TORCH2NP = {torch.int32: np.dtype(np.int32), torch.int64: np.dtype(np.int64), torch.float32: np.dtype(np.float32),
             torch.float64: np.dtype(np.float64)}


def _check(err):
    if err.str is not None:
        raise RuntimeError(err.str.decode())


def _fn(name, argtypes):
    f = getattr(lib(), name)
    f.argtypes = argtypes
    f.restype = KernelError
    return f


def min_identity(dtype: np.dtype):
    return int(np.iinfo(dtype).max) if dtype.kind in "iu" else float("inf")


class _Group:
    def __init__(self, dtype, members, arrays, device, variant):
        self.dtype = dtype
        self.name = _NAME[dtype]
        self.members = members
        self.variant = variant
        self.counter = torch.zeros(2, dtype=torch.int64, device=device)
        self.outputs = {}
        self._build(arrays, device)

    def _build(self, arrays, device):
        lib_ = lib()
        tile_lists = lib_.awkward_hip_batch_reduce_tile_lists
        tile_lists.argtypes = [ctypes.c_int]
        tile_lists.restype = ctypes.c_int64
        L = int(tile_lists(self.variant))
        sort_lists = int(lib_.awkward_hip_batch_sort_tile_lists())
        red_rows, sort_rows = [], []
        rtile = stile = elem = 0
        for i in self.members:
            offsets, content = arrays[i]
            n = offsets.shape[0] - 1
            out = {
                "sum": torch.empty(n, dtype=_SUM[self.dtype], device=device),
                "min": torch.empty(n, dtype=content.dtype, device=device),
                "argmin": torch.empty(n, dtype=torch.int64, device=device),
                "mask": torch.empty(n, dtype=torch.int8, device=device),
                "argsort": torch.empty(max(content.shape[0], 1), dtype=torch.int64, device=device),
            }
            self.outputs[i] = out
            red_rows.append([content.data_ptr(), offsets.data_ptr(), out["sum"].data_ptr(), out["min"].data_ptr(),
                             out["argmin"].data_ptr(), out["mask"].data_ptr(), n, rtile])
            sort_rows.append([content.data_ptr(), offsets.data_ptr(), out["argsort"].data_ptr(), n,
                              content.shape[0], stile, elem])
            rtile += -(-n // L) if n > 0 else 0
            stile += -(-n // sort_lists) if n > 0 else 0
            elem += content.shape[0]
        self.reduce_tiles, self.sort_tiles, self.elements = rtile, stile, elem
        self.red_table = torch.from_numpy(np.array(red_rows, dtype=np.uint64).view(np.int64)).to(device)
        self.sort_table = torch.from_numpy(np.array(sort_rows, dtype=np.uint64).view(np.int64)).to(device)
        P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        self._reset = _fn("awkward_hip_batch_reset", [P])
        self._reduce = _fn(f"awkward_hip_batch_reduce_{self.name}", [P, I32, I64, P, _CT[self.name], I32])
        self._sort = _fn(f"awkward_hip_batch_argsort_{self.name}", [P, I32, I64, P, ctypes.c_bool, I64])
        self.identity = min_identity(self.dtype)

    def run_reduce(self):
        n = len(self.members)
        _check(self._reset(self.counter.data_ptr()))
        _check(self._reduce(self.red_table.data_ptr(), n, self.reduce_tiles, self.counter.data_ptr(),
                            self.identity, self.variant))

    def run_sort(self):
        n = len(self.members)
        c = self.counter.data_ptr() + 8
        _check(self._reset(c))
        _check(self._sort(self.sort_table.data_ptr(), n, self.sort_tiles, c, True, self.elements))


class BatchRunner:
    def __init__(self, arrays, variants=None, argsort=True):
        """arrays: list of (offsets, content) torch tensors on one device."""
        self.arrays = arrays
        self.argsort = argsort
        device = arrays[0][0].device
        groups = {}
        for i, (offsets, content) in enumerate(arrays):
            groups.setdefault(_TORCH2NP[content.dtype], []).append(i)
        variants = variants or {}
        self.groups = [_Group(dt, members, arrays, device, variants.get(_NAME[dt], 0))
                       for dt, members in sorted(groups.items(), key=lambda kv: kv[0].name)]
        # sort scratch sized before any capture: the largest group
        need = max(g.elements for g in self.groups)
        reserve = lib().awkward_hip_reserve_scratch
        reserve(int(24 * max(need, 1) + 4096))

    def run(self):
        set_stream()
        for g in self.groups:
            g.run_reduce()
            if self.argsort:
                g.run_sort()
        return self

    def result(self, i):
        for g in self.groups:
            if i in g.outputs:
                out = dict(g.outputs[i])
                out["argsort"] = out["argsort"][: self.arrays[i][1].shape[0]]
                return out
        raise KeyError(i)

    def algorithmic_bytes(self):
        """Compulsory traffic of one run(): reduce reads offsets + content and
        writes four outputs per list; argsort reads offsets + content and
        writes one int64 per element."""
        total = 0
        for offsets, content in self.arrays:
            n = offsets.shape[0] - 1
            N = content.shape[0]
            s = content.element_size()
            sum_s = 8 if s == 4 and content.dtype == torch.int32 else s
            total += 8 * (n + 1) + s * N + n * (sum_s + s + 8 + 1)
            if self.argsort:
                total += 8 * (n + 1) + s * N + 8 * N
        return total

    def tune(self, reps=3):
        """Time every reduce variant per dtype group on this batch; keep the
        fastest (the autotuner's choice) and return the timings."""
        timings = {}
        for g in self.groups:
            best, best_ms = g.variant, None
            times = {}
            for v in range(len(REDUCE_VARIANTS)):
                g.variant = v
                g._build(self.arrays, self.arrays[0][0].device)
                g.run_reduce()
                torch.cuda.synchronize()
                samples = []
                for _ in range(reps):
                    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    s.record()
                    g.run_reduce()
                    e.record()
                    torch.cuda.synchronize()
                    samples.append(s.elapsed_time(e))
                ms = sorted(samples)[len(samples) // 2]
                times[v] = ms
                if best_ms is None or ms < best_ms:
                    best, best_ms = v, ms
            g.variant = best
            g._build(self.arrays, self.arrays[0][0].device)
            timings[g.name] = {"best": best, "ms": times}
        return timings
