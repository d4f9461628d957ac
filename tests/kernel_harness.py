"""Runs the reference's golden kernel vectors against a kernel library.

A vector (from ``tests/golden/kernel_vectors.json``) names its inputs and
expected outputs by argument name, without types.  Like the reference's
generated kernel tests (``dev/generate-tests.py:530-662``) it is run against
every specialization of its family whose argument types can represent the
vector's values; output buffers are pre-filled with the sentinel 123.

The memory side is pluggable so the same vectors drive the CPU oracle (numpy
host buffers) and the hip library (torch device buffers).
"""

from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

from awkward_amd._kernel_signatures import SPECIALIZATIONS

HERE = os.path.dirname(os.path.abspath(__file__))

NP_TYPES = {
    "bool": np.bool_,
    "int8": np.int8,
    "uint8": np.uint8,
    "int16": np.int16,
    "uint16": np.uint16,
    "int32": np.int32,
    "uint32": np.uint32,
    "int64": np.int64,
    "uint64": np.uint64,
    "float": np.float32,
    "double": np.float64,
}


def load_vectors():
    with open(os.path.join(HERE, "golden", "kernel_vectors.json")) as f:
        return json.load(f)


def parse(spec_type: str):
    t = spec_type
    if t.startswith("Const["):
        t = t[6:-1]
    depth = 0
    while t.startswith("List["):
        depth += 1
        t = t[5:-1]
    if t.endswith("_t"):
        t = t[:-2]
    return depth, NP_TYPES[t]


def fits(values, dtype) -> bool:
    dtype = np.dtype(dtype)
    flat = []

    def walk(v):
        if isinstance(v, list):
            for x in v:
                walk(x)
        else:
            flat.append(v)

    walk(values)
    for v in flat:
        if isinstance(v, float) and (math.isnan(v) or math.isinf(v)):
            if dtype.kind != "f":
                return False
            continue
        if dtype.kind == "b":
            continue  # ctypes.c_bool truthifies, as the reference's tests do
        elif dtype.kind in "iu":
            if float(v) != int(v):
                return False
            info = np.iinfo(dtype)
            if not info.min <= int(v) <= info.max:
                return False
    return True


def specializations(family: str):
    return [n for n, (fam, _, _) in SPECIALIZATIONS.items() if fam == family]


class Case:
    """One (specialization, vector) pair, laid out as typed numpy buffers."""

    def __init__(self, name, vector):
        self.name = name
        self.vector = vector
        self.args = []  # (kind, value): kind in {"ptr", "ptrs", "scalar"}
        self.outputs = {}  # argname -> index into self.args
        self.ok = True
        _, args, _ = SPECIALIZATIONS[name]
        inputs, outputs = vector["inputs"], vector.get("outputs", {})
        for argname, spec_type, direction in args:
            depth, dtype = parse(spec_type)
            if depth == 0:
                value = inputs[argname]
                if not fits([value], dtype):
                    self.ok = False
                self.args.append(("scalar", value))
                continue
            if argname in inputs and (depth == 1):
                values = inputs[argname]
                if not fits(values, dtype):
                    self.ok = False
                    values = []
                buf = np.array(values, dtype=dtype) if values else np.zeros(0, dtype)
                if argname in outputs:
                    want = max(len(outputs[argname]), len(buf), 1)
                    full = np.full(want, 123, dtype=dtype)
                    full[: len(buf)] = buf
                    buf = full
                elif len(buf) == 0:
                    buf = np.zeros(1, dtype)  # never read: keeps pointers non-null
                self.args.append(("ptr", buf))
            elif depth == 1:
                expect = outputs.get(argname, [])
                if not fits(expect, dtype):
                    self.ok = False
                self.args.append(("ptr", np.full(max(len(expect), 1), 123, dtype=dtype)))
            else:  # List[List[T]] output (combinations' tocarry)
                expect = outputs[argname]
                if not fits(expect, dtype):
                    self.ok = False
                bufs = [np.full(max(len(e), 1), 123, dtype=dtype) for e in expect]
                self.args.append(("ptrs", bufs))
            if argname in outputs:
                self.outputs[argname] = len(self.args) - 1

    def expected(self, argname):
        return self.vector["outputs"][argname]


def compare(name, argname, got, expect):
    got = np.asarray(got)
    if got.dtype.kind == "f":
        exp = np.asarray(expect, dtype=np.float64)
        assert np.allclose(got[: len(exp)].astype(np.float64), exp, equal_nan=True), (
            name, argname, got[: len(exp)].tolist(), expect)
    else:
        assert got[: len(expect)].tolist() == [type(g)(e) for g, e in zip(got.tolist(), expect)], (
            name, argname, got[: len(expect)].tolist(), expect)


class HostMemory:
    """Buffers stay numpy arrays; pointers are their host addresses."""

    def prepare(self, case: Case):
        keep, call = [], []
        for kind, v in case.args:
            if kind == "scalar":
                call.append(v)
            elif kind == "ptr":
                keep.append(v)
                call.append(v.ctypes.data)
            else:
                arr = (ctypes.c_void_p * len(v))(*[b.ctypes.data for b in v])
                keep.append(arr)
                call.append(ctypes.addressof(arr))
        return keep, call

    def fetch(self, case: Case, keep, index):
        kind, v = case.args[index]
        return v if kind == "ptr" else v


def run_case(fn, case: Case, memory):
    keep, call = memory.prepare(case)
    err = fn(*call)
    return err, keep


def check_case(fn, case: Case, memory):
    err, keep = run_case(fn, case, memory)
    if case.vector.get("error"):
        assert err.str is not None, (case.name, "expected error", case.vector["message"])
        assert err.str.decode() == case.vector["message"], (case.name, err.str)
        return
    assert err.str is None, (case.name, err.str)
    for argname, index in case.outputs.items():
        kind, _ = case.args[index]
        got = memory.fetch(case, keep, index)
        if kind == "ptr":
            compare(case.name, argname, got, case.expected(argname))
        else:
            for g, e in zip(got, case.expected(argname)):
                compare(case.name, argname, g, e)


class DeviceMemory:
    """Buffers copied to torch tensors on cuda:0; pointers are device addresses."""

    def prepare(self, case: Case):
        import torch

        from awkward_amd._kernels import PointerArray

        keep, call = [], []
        for kind, v in case.args:
            if kind == "scalar":
                call.append(v)
            elif kind == "ptr":
                t = torch.from_numpy(np.ascontiguousarray(v)).to("cuda")
                keep.append(t)
                call.append(t.data_ptr())
            else:
                ts = [torch.from_numpy(np.ascontiguousarray(b)).to("cuda") for b in v]
                table = PointerArray(ts)
                keep.append(table)
                call.append(table.data_ptr())
        return keep, call

    def fetch(self, case: Case, keep, index):
        import torch

        torch.cuda.synchronize()
        kind, _ = case.args[index]
        # keep holds one entry per non-scalar argument, in order
        pos = sum(1 for k, _ in case.args[:index] if k != "scalar")
        item = keep[pos]
        if kind == "ptr":
            return item.cpu().numpy()
        return [t.cpu().numpy() for t in item.tensors]
