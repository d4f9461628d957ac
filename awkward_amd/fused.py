"""Fused, synchronisation-free forms of the configs' op sequences, and
hipGraph capture of them.

``EventPipeline`` evaluates, for every event of a list-of-records array,

    sel = ev[(term_1) & (term_2) & ...][:, start:stop:step]
    ak.num(sel), ak.sum(sel[field], axis=-1), ak.argmax(sel[field], axis=-1), ak.max(sel[field], axis=-1)

in one kernel (``awkward_hip_event_pipeline``) with the same results as the
per-kernel path; ``pair_max`` evaluates ak.max / ak.argmax over
``ak.combinations(ev, 2)`` of the pair expression
``2*pt1*pt2*(cosh(eta1-eta2) - cos(phi1-phi2))`` without materialising the
pairs (``awkward_hip_pair_max``), and ``PairReduce`` does the same for any
elementwise expression of the pairs' fields, given as a program
(``awkward_hip_pair_reduce``).  Neither reads anything back to the host,
so ``Graph`` can capture them (with torch.cuda.CUDAGraph, which is a hipGraph
on ROCm) and replay the whole sequence with one launch.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from awkward_amd import contents as C
from awkward_amd._kernels import KernelError, device_address, set_stream
from awkward_amd._lib import lib

_OPS = {">": 0, ">=": 1, "<": 2, "<=": 3, "==": 4, "!=": 5}
_TYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.int8): 2, np.dtype(np.int32): 3,
          np.dtype(np.int64): 4}
_NONE = 2**63 - 1

_fn_cache = {}


def _fn(name, argtypes):
    f = _fn_cache.get(name)
    if f is None:
        f = getattr(lib(), name)
        f.argtypes = argtypes
        f.restype = KernelError
        _fn_cache[name] = f
    return f


def _check(err):
    if err.str is not None:
        raise RuntimeError(err.str.decode())


def _records(array):
    layout = array.layout
    if not isinstance(layout, C.ListOffsetArray) or not isinstance(layout.content, C.RecordArray):
        raise NotImplementedError("fused pipelines take a list array of records")
    return layout.offsets, layout.content


class EventPipeline:
    """terms: [(field, op, value) or (field, "abs"+op, value)], combined with &."""

    def __init__(self, events, terms, slice_=slice(None), reduce_field="pt",
                 outputs=("num", "sum", "argmax", "max")):
        offsets, record = _records(events)
        self.offsets = offsets
        self.n = record.backend.nplike.length(offsets) - 1
        names = []
        for f, _, _ in terms:
            if f not in names:
                names.append(f)
        if reduce_field not in names:
            names.append(reduce_field)
        cols = [record._getitem_field(f).data for f in names]
        self.cols = cols
        nplike = record.backend.nplike
        self.n_particles = nplike.length(cols[0])
        self.field_ptrs = (ctypes.c_void_p * len(cols))(*[c.data_ptr() for c in cols])
        self.field_types = (ctypes.c_int32 * len(cols))(*[_TYPES[nplike.dtype(c)] for c in cols])
        tf, to, ta, tv = [], [], [], []
        for f, op, value in terms:
            absolute = op.startswith("abs")
            op = op[3:] if absolute else op
            tf.append(names.index(f))
            to.append(_OPS[op])
            ta.append(1 if absolute else 0)
            tv.append(float(value))
        k = len(terms)
        self.terms = ((ctypes.c_int32 * k)(*tf), (ctypes.c_int32 * k)(*to), (ctypes.c_int32 * k)(*ta),
                      (ctypes.c_double * k)(*tv), k)
        self.start = _NONE if slice_.start is None else slice_.start
        self.stop = _NONE if slice_.stop is None else slice_.stop
        self.step = 1 if slice_.step is None else slice_.step
        self.red = names.index(reduce_field)
        dev = offsets.device
        spec = {"num": [("num", torch.int64)], "sum": [("sum", torch.float32)],
                "argmax": [("argmax", torch.int64), ("argmax_mask", torch.int8)],
                "max": [("max", torch.float32), ("max_mask", torch.int8)]}
        self.out = {}
        for name in outputs:
            for key, dt in spec[name]:
                self.out[key] = torch.empty(self.n, dtype=dt, device=dev)
        P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        self._f = _fn("awkward_hip_event_pipeline",
                      [P, I64, I64, P, P, I32, P, P, P, P, I32, I64, I64, I64, I32, P, P, P, P, P, P])

    def run(self):
        set_stream()
        tf, to, ta, tv, k = self.terms
        o = self.out
        _check(self._f(device_address(self.offsets), self.n, self.n_particles, ctypes.addressof(self.field_ptrs),
                       ctypes.addressof(self.field_types), len(self.cols), ctypes.addressof(tf),
                       ctypes.addressof(to), ctypes.addressof(ta), ctypes.addressof(tv), k, self.start, self.stop,
                       self.step, self.red, *[o[k].data_ptr() if k in o else 0 for k in
                                              ("num", "sum", "argmax", "argmax_mask", "max", "max_mask")]))
        return o

    def algorithmic_bytes(self, n_particles: int) -> int:
        """Compulsory traffic: offsets and the used fields read once, every
        requested output written once."""
        n = self.n
        read = 8 * (n + 1) + sum(c.element_size() * n_particles for c in self.cols)
        return read + sum(v.element_size() * n for v in self.out.values())


class PairMax:
    """ak.max / ak.argmax over ak.combinations(ev, 2) of the pair expression."""

    def __init__(self, events, pt="pt", eta="eta", phi="phi"):
        offsets, record = _records(events)
        self.offsets = offsets
        self.n = record.backend.nplike.length(offsets) - 1
        self.cols = [record._getitem_field(f).data for f in (pt, eta, phi)]
        dev = offsets.device
        self.out = {
            "max": torch.empty(self.n, dtype=torch.float32, device=dev),
            "argmax": torch.empty(self.n, dtype=torch.int64, device=dev),
            "mask": torch.empty(self.n, dtype=torch.int8, device=dev),
        }
        P, I64 = ctypes.c_void_p, ctypes.c_int64
        self._f = _fn("awkward_hip_pair_max", [P, I64, P, P, P, P, P, P, ctypes.c_float])

    def run(self):
        set_stream()
        o = self.out
        _check(self._f(device_address(self.offsets), self.n, *[c.data_ptr() for c in self.cols], o["max"].data_ptr(),
                       o["argmax"].data_ptr(), o["mask"].data_ptr(), float("-inf")))
        return o


class PairReduce:
    """ak.max / ak.min / ak.argmax / ak.argmin / ak.sum (axis=-1) of one
    elementwise expression over the pairs of ``ak.combinations(ev, 2)``,
    from a program whose leaves read leg 0 or leg 1 of each pair
    (``awkward_hip_pair_reduce``): the pairs and the expression's values are
    never written.  ``expr`` is an ``elementwise.DeviceExpr`` of float32 or
    float64; ``offsets`` are the lists' int64 offsets into its columns."""

    _SLOTS = ("max", "min", "sum", "argmax", "argmin", "mask")

    def __init__(self, offsets, expr, outputs=("max", "argmax")):
        self.offsets = offsets
        self.n = offsets.shape[0] - 1
        self.program, self._keep = expr.program()
        dt = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}.get(np.dtype(expr.dtype))
        if dt is None:
            raise NotImplementedError(f"pair reductions of {expr.dtype} expressions")
        dev = offsets.device
        self.out = {}
        for name in outputs:
            if name not in self._SLOTS[:5]:
                raise ValueError(f"unknown pair reduction {name!r}")
            self.out[name] = torch.empty(self.n, dtype=dt if name in ("max", "min", "sum") else torch.int64,
                                         device=dev)
        if any(o != "sum" for o in outputs):
            self.out["mask"] = torch.empty(self.n, dtype=torch.int8, device=dev)
        P, I64 = ctypes.c_void_p, ctypes.c_int64
        self._f = _fn("awkward_hip_pair_reduce", [P, I64, P, P, P, P, P, P, P])

    def run(self):
        set_stream()
        o = self.out
        _check(self._f(device_address(self.offsets), self.n, ctypes.addressof(self.program),
                       *[o[k].data_ptr() if k in o else 0 for k in self._SLOTS]))
        return o


class _Tuples:
    """One n's outputs of a CombinationsPairs launch (the shape of
    Combinations's: tooffsets, total, carry)."""

    def __init__(self, tooffsets, totallen):
        self.tooffsets, self.totallen = tooffsets, totallen
        self.total, self.carry = 0, []


class CombinationsPairs:
    """ak.combinations(ev, n) for up to two n in {2, 3} (no replacement) and,
    optionally, the pair reductions of one expression (as PairReduce), in one
    launch (``awkward_hip_combinations_pairs``) after one
    combinations_length scan per n: each workgroup writes its lists' carries
    (write-bound) and evaluates and reduces the program over their pairs
    (ALU-bound), so the two overlap on the chip.  The tuple counts are read
    once at construction to size the carries; run() is capturable."""

    def __init__(self, events, ns, expr=None, outputs=("max", "argmax")):
        layout = events.layout
        self.offsets = layout.offsets
        self.length = layout.length
        self.ns = tuple(ns)
        if not 1 <= len(self.ns) <= 2 or any(n not in (2, 3) for n in self.ns) or len(set(self.ns)) != len(self.ns):
            raise ValueError("one or two distinct tuple sizes from {2, 3}")
        dev = self.offsets.device
        self.starts, self.stops = self.offsets[:-1], self.offsets[1:]
        self.tuples = {n: _Tuples(torch.empty(self.length + 1, dtype=torch.int64, device=dev),
                                  torch.empty(1, dtype=torch.int64, device=dev)) for n in self.ns}
        P, I64, I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        self._len = _fn("awkward_ListArray64_combinations_length_64", [P, P, I64, ctypes.c_bool, P, P, I64])
        self._f = _fn("awkward_hip_combinations_pairs", [P, I64, I32, P, P, P, P, P, P, P, P, P, P])
        self._run_lengths()
        torch.cuda.synchronize()
        carries = []
        for n in self.ns:
            t = self.tuples[n]
            t.total = int(t.totallen.item())
            t.carry = [torch.empty(max(t.total, 1), dtype=torch.int64, device=dev) for _ in range(n)]
            carries.extend(t.carry)
        k = len(self.ns)
        self.req_n = (ctypes.c_int32 * k)(*self.ns)
        self.req_tooffsets = (ctypes.c_void_p * k)(*[self.tuples[n].tooffsets.data_ptr() for n in self.ns])
        self.req_carries = (ctypes.c_void_p * len(carries))(*[c.data_ptr() for c in carries])
        self.out = {}
        self.program = None
        if expr is not None:
            pr = PairReduce(self.offsets, expr, outputs)
            self.program, self._keep, self.out = pr.program, pr._keep, pr.out

    def _run_lengths(self):
        set_stream()
        for n in self.ns:
            t = self.tuples[n]
            _check(self._len(t.totallen.data_ptr(), t.tooffsets.data_ptr(), n, False, self.starts.data_ptr(),
                             self.stops.data_ptr(), self.length))

    def run(self):
        self._run_lengths()
        o = self.out
        _check(self._f(device_address(self.offsets), self.length, len(self.ns), ctypes.addressof(self.req_n),
                       ctypes.addressof(self.req_tooffsets), ctypes.addressof(self.req_carries),
                       ctypes.addressof(self.program) if self.program is not None else 0,
                       *[o[k].data_ptr() if k in o else 0 for k in PairReduce._SLOTS]))
        return self

    def algorithmic_bytes(self):
        """Offsets read once (and once per scan), tooffsets written, every
        carry written once."""
        b = 8 * (self.length + 1) * (1 + 2 * len(self.ns))
        return b + sum(8 * n * self.tuples[n].total for n in self.ns)


class Combinations:
    """ak.combinations(ev, n)'s carries without a host round trip per call:
    the tuple count is read once at construction to size the outputs, then
    run() enqueues combinations_length + the decode (capturable)."""

    def __init__(self, events, n, replacement=False):
        layout = events.layout
        self.offsets = layout.offsets
        self.length = layout.length
        self.n, self.replacement = n, replacement
        dev = self.offsets.device
        self.tooffsets = torch.empty(self.length + 1, dtype=torch.int64, device=dev)
        self.totallen = torch.empty(1, dtype=torch.int64, device=dev)
        self.starts = self.offsets[:-1]
        self.stops = self.offsets[1:]
        P, I64 = ctypes.c_void_p, ctypes.c_int64
        self._len = _fn("awkward_ListArray64_combinations_length_64", [P, P, I64, ctypes.c_bool, P, P, I64])
        self._dec = _fn("awkward_hip_combinations_from_offsets", [P, P, P, P, P, I64, I64, ctypes.c_bool, I64])
        self._run_length()
        torch.cuda.synchronize()
        self.total = int(self.totallen.item())
        self.carry = [torch.empty(max(self.total, 1), dtype=torch.int64, device=dev) for _ in range(n)]
        from awkward_amd._kernels import PointerArray

        self.table = PointerArray(self.carry)
        self.toindex = torch.empty(n, dtype=torch.int64, device=dev)

    def _run_length(self):
        set_stream()
        _check(self._len(self.totallen.data_ptr(), self.tooffsets.data_ptr(), self.n, self.replacement,
                         self.starts.data_ptr(), self.stops.data_ptr(), self.length))

    def run(self):
        self._run_length()
        _check(self._dec(self.table.data_ptr(), self.toindex.data_ptr(), self.tooffsets.data_ptr(),
                         self.starts.data_ptr(), self.stops.data_ptr(), self.length, self.n, self.replacement,
                         self.total))
        return self.carry

    def algorithmic_bytes(self):
        return 16 * self.length + 8 * (self.length + 1) + 8 * self.n * self.total


def pair_expr(p0, p1):
    """The pair expression over gathered legs (float32 device tensors)."""
    out = torch.empty_like(p0[0])
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    f = _fn("awkward_hip_pair_expr", [P, P, P, P, P, P, P, I64])
    set_stream()
    _check(f(out.data_ptr(), *[t.data_ptr() for t in p0], *[t.data_ptr() for t in p1], out.shape[0]))
    return out


class Concurrent:
    """Runs independent op sequences side by side: each on its own stream
    (default priority) forked from the current one, joined back before
    returning.  Under capture the branches become parallel branches of the
    hipGraph, so a latency-bound kernel overlaps a bandwidth-bound one.
    The sequences must not share outputs."""

    def __init__(self, *fns):
        self.fns = fns
        self.streams = [torch.cuda.Stream() for _ in fns[1:]]

    def __call__(self):
        cur = torch.cuda.current_stream()
        results = [None] * len(self.fns)
        for st, fn, k in zip(self.streams, self.fns[1:], range(1, len(self.fns))):
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                results[k] = fn()
        results[0] = self.fns[0]()
        for st in self.streams:
            cur.wait_stream(st)
        return tuple(results)


class Graph:
    """Captures fn() (which must not synchronise or allocate through the
    library) into a hipGraph and replays it.

    The warm-up runs and the capture use one stream, so the library's
    per-stream state and scratch pools are allocated (at the sizes fn needs)
    before the capture begins; the captured kernels keep using that stream's
    pools, which are never freed while the process runs."""

    def __init__(self, fn, warmup=2):
        self.fn = fn
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.stream = s
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=s):
            self.result = fn()
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()
        return self.result
