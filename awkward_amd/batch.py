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

``WaveRunner`` streams a batch larger than HBM through ``BatchRunner`` in
waves: consecutive arrays whose device bytes (``device_bytes``) fit a budget
are made on the device, run and dropped before the next wave.

``variants`` selects the reduce tile per dtype (the ten tiles of
``dispatch.VARIANTS``); the autotuner (``tune()``) times them on the batch
itself and keeps the fastest.  ``variants="learned"`` instead enqueues, per
dtype group, the dispatcher's launch (a statistics pass over all the group's
offsets whose last workgroup runs the MLP inference) and the ten tile
launches predicated on its choice
(``awkward_hip_batch_reduce_<T>_chosen``) -- still no host synchronisation.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from awkward_amd import dispatch
from awkward_amd._kernels import KernelError, set_stream
from awkward_amd._lib import lib

#: (threads, lists per thread, LDS bytes) of awkward_hip_batch_reduce_* variants
REDUCE_VARIANTS = dispatch.VARIANTS

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
    def __init__(self, dtype, members, arrays, device, variant, dispatcher=None):
        self.dtype = dtype
        self.name = _NAME[dtype]
        self.members = members
        self.variant = variant
        self.dispatcher = dispatcher
        self.counter = torch.zeros(2, dtype=torch.int64, device=device)
        self.outputs = {}
        self._build(arrays, device)

    def _reduce_table(self, arrays, L, device):
        rows, tile = [], 0
        for i in self.members:
            offsets, content = arrays[i]
            n = offsets.shape[0] - 1
            out = self.outputs[i]
            rows.append([content.data_ptr(), offsets.data_ptr(), out["sum"].data_ptr(), out["min"].data_ptr(),
                         out["argmin"].data_ptr(), out["mask"].data_ptr(), n, tile])
            tile += -(-n // L) if n > 0 else 0
        return torch.from_numpy(np.array(rows, dtype=np.uint64).view(np.int64)).to(device), tile

    def _build(self, arrays, device):
        lib_ = lib()
        tile_lists = lib_.awkward_hip_batch_reduce_tile_lists
        tile_lists.argtypes = [ctypes.c_int]
        tile_lists.restype = ctypes.c_int64
        sort_lists = int(lib_.awkward_hip_batch_sort_tile_lists())
        sort_rows = []
        stile = elem = 0
        for i in self.members:
            offsets, content = arrays[i]
            n = offsets.shape[0] - 1
            if i not in self.outputs:
                self.outputs[i] = {
                    "sum": torch.empty(n, dtype=_SUM[self.dtype], device=device),
                    "min": torch.empty(n, dtype=content.dtype, device=device),
                    "argmin": torch.empty(n, dtype=torch.int64, device=device),
                    "mask": torch.empty(n, dtype=torch.int8, device=device),
                    "argsort": torch.empty(max(content.shape[0], 1), dtype=torch.int64, device=device),
                }
            out = self.outputs[i]
            sort_rows.append([content.data_ptr(), offsets.data_ptr(), out["argsort"].data_ptr(), n,
                              content.shape[0], stile, elem])
            stile += -(-n // sort_lists) if n > 0 else 0
            elem += content.shape[0]
        self.sort_tiles, self.elements = stile, elem
        self.sort_table = torch.from_numpy(np.array(sort_rows, dtype=np.uint64).view(np.int64)).to(device)
        P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        self._reset = _fn("awkward_hip_batch_reset", [P])
        self._reduce = _fn(f"awkward_hip_batch_reduce_{self.name}", [P, I32, I64, P, _CT[self.name], I32, I64])
        self._sort = _fn(f"awkward_hip_batch_argsort_{self.name}", [P, I32, I64, P, ctypes.c_bool, I64])
        self.identity = min_identity(self.dtype)
        if self.dispatcher is None:
            self.red_table, self.reduce_tiles = self._reduce_table(arrays, int(tile_lists(self.variant)), device)
            return
        # learned: a table per tile shape, the statistics plan of the whole group
        self._chosen = _fn(f"awkward_hip_batch_reduce_{self.name}_chosen", [P, P, I32, P, _CT[self.name], P, I64])
        self.var_tables = [self._reduce_table(arrays, int(tile_lists(v)), device) for v in range(len(REDUCE_VARIANTS))]
        self._ptrs = (ctypes.c_void_p * len(REDUCE_VARIANTS))(*[t.data_ptr() for t, _ in self.var_tables])
        self._tiles = (ctypes.c_int64 * len(REDUCE_VARIANTS))(*[n for _, n in self.var_tables])
        entries = [(arrays[i][0], arrays[i][0].shape[0] - 1, 0) for i in self.members]
        sizes = [(sum(arrays[i][0].shape[0] - 1 for i in self.members),
                  sum(arrays[i][1].shape[0] for i in self.members))]
        from awkward_amd.learned import OP_CODE

        self.plan, self.meta, self.choice = self.dispatcher.choose_device(
            OP_CODE["batch"], entries, sizes, [self.dtype.itemsize], device)

    def run_reduce(self):
        n = len(self.members)
        _check(self._reset(self.counter.data_ptr()))
        if self.dispatcher is None:
            _check(self._reduce(self.red_table.data_ptr(), n, self.reduce_tiles, self.counter.data_ptr(),
                                self.identity, self.variant, self.elements))
            return
        self.dispatcher.run_device(self.plan, self.meta, self.choice)
        _check(self._chosen(ctypes.addressof(self._ptrs), ctypes.addressof(self._tiles), n, self.counter.data_ptr(),
                            self.identity, self.choice.data_ptr(), self.elements))

    def chosen(self):
        """The learned tile (reads the device choice: synchronises)."""
        return int(self.choice[0].item()) if self.dispatcher is not None else self.variant

    def run_sort(self):
        n = len(self.members)
        c = self.counter.data_ptr() + 8
        _check(self._reset(c))
        _check(self._sort(self.sort_table.data_ptr(), n, self.sort_tiles, c, True, self.elements))


class BatchRunner:
    def __init__(self, arrays, variants=None, argsort=True, dispatcher=None):
        """arrays: list of (offsets, content) torch tensors on one device;
        variants: {dtype name: tile} or "learned" (then `dispatcher`, or the
        default trained one, chooses on the device)."""
        self.arrays = arrays
        self.argsort = argsort
        device = arrays[0][0].device
        groups = {}
        for i, (offsets, content) in enumerate(arrays):
            groups.setdefault(_TORCH2NP[content.dtype], []).append(i)
        if variants == "learned":
            if dispatcher is None:
                from awkward_amd.learned import Dispatcher

                dispatcher = Dispatcher()
            variants = {}
        else:
            dispatcher = None
        variants = variants or {}
        self.groups = [_Group(dt, members, arrays, device, variants.get(_NAME[dt], 0), dispatcher)
                       for dt, members in sorted(groups.items(), key=lambda kv: kv[0].name)]
        # scratch and the reductions' queue sized before any capture: the
        # largest group (the sort's merge buffers and chunk queue; the
        # reductions' long-list chunks and partials)
        need = max(max(g.elements for g in self.groups), 1)
        lib_ = lib()
        lib_.awkward_hip_reserve_scratch.argtypes = [ctypes.c_int64]
        lib_.awkward_hip_reserve_queue.argtypes = [ctypes.c_int64]
        sort_bytes = 24 * need + (need // 65 + 2) * 48 + (need // 4096 + need // 1025 + 4) * 64 + 8192
        set_stream()  # the pools belong to the current stream
        if lib_.awkward_hip_reserve_scratch(int(sort_bytes)):
            raise MemoryError("cannot reserve the batch runner's scratch memory")
        if lib_.awkward_hip_reserve_queue(int((need // 8192 + need // 4097 + 4) * 96 + 8192)):
            raise MemoryError("cannot reserve the batch runner's queue memory")

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
            if g.dispatcher is not None:
                raise ValueError("tune() times fixed tiles; this runner chooses them with the learned dispatcher")
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


def release_pools():
    """Frees the current stream's scratch and queue pools (after synchronising
    it; the next call allocates afresh): call only when no captured hipGraph
    still uses them."""
    f = lib().awkward_hip_release_pools
    f.restype = ctypes.c_int
    set_stream()
    if f() != 0:
        raise RuntimeError("awkward_hip_release_pools failed")


def device_bytes(n_lists: int, n_elements: int, itemsize: int, argsort: bool = True) -> int:
    """HBM that one array holds while a BatchRunner runs it: offsets,
    content, the four reduce outputs (the sum counted at 8 bytes, its widest
    type), and (argsort) its int64 output plus its share of the sort's
    scratch (24 bytes per element of the dtype group)."""
    b = 8 * (n_lists + 1) + itemsize * n_elements + n_lists * (8 + itemsize + 8 + 1)
    if argsort:
        b += 8 * n_elements + 24 * n_elements
    return b


def plan_waves(sizes, budget: int):
    """Consecutive runs of array indices whose bytes stay within `budget`
    (an array larger than the budget is a wave of its own)."""
    waves, cur, used = [], [], 0
    for i, b in enumerate(sizes):
        if cur and used + b > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += b
    if cur:
        waves.append(cur)
    return waves


class WaveRunner:
    """A batch streamed through the device in waves.

    ``specs``: per array (n_lists, n_elements, numpy dtype); ``make(i)``
    returns array i's (offsets, content) device tensors (copied or generated
    on demand).  Iterating yields ``(indices, arrays)`` per wave -- run them
    (a BatchRunner over ``arrays``, or the registry path) and read the
    results inside the loop body: the wave's tensors and the library's
    scratch pools are released before the next wave is made (so hipGraphs
    captured over a wave must not outlive it)."""

    def __init__(self, specs, make, budget_bytes: int, argsort: bool = True):
        self.specs = list(specs)
        self.make = make
        self.sizes = [device_bytes(n, N, np.dtype(dt).itemsize, argsort) for n, N, dt in self.specs]
        self.waves = plan_waves(self.sizes, budget_bytes)

    def __iter__(self):
        for wave in self.waves:
            arrays = [self.make(i) for i in wave]
            yield wave, arrays
            del arrays
            release_pools()
            torch.cuda.empty_cache()

    def bytes(self, wave):
        return sum(self.sizes[i] for i in wave)
