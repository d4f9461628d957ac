"""Autotuner for the segmented reductions' tile variants.

``time_variants`` times every compiled variant of one reduction on the
device (median of ``reps`` launches, each after an L2/MALL scrub so the
inputs come from HBM).  The scrub only reads (a sum over 1 GiB): a scrub
that writes leaves up to 256 MiB of dirty lines in the caches, whose
write-back would be charged to the timed kernel.  ``Autotuner`` keeps the fastest variant per
(op, dtype, list-count bucket, mean-length bucket) key in a JSON table;
``install()`` makes the hip backend's reducers consult it (falling back to
the heuristic of ``awkward_amd.dispatch`` for keys it has not measured).
"""

from __future__ import annotations

import json
import math

import numpy as np

from awkward_amd import dispatch


def bucket(n_lists: int, n_elements: int):
    mean = n_elements / max(n_lists, 1)
    return (int(round(math.log2(max(n_lists, 1)))), int(round(2 * math.log2(mean + 1))))


def time_variants(op, out, data, offsets, n, mask=None, identity=None, reps=5, scrub=None, variants=None):
    import torch

    variants = range(len(dispatch.VARIANTS)) if variants is None else variants
    if scrub is None:
        scrub = torch.empty(1 << 28, dtype=torch.float32, device=data.device)

    def call(v):
        if op == "sum":
            return dispatch.reduce_sum(out, data, offsets, n, v)
        if op in ("max", "min"):
            return dispatch.reduce_extremum(op, out, mask, data, offsets, n, identity, v)
        return dispatch.reduce_arg(op, out, mask, data, offsets, n, v)

    times = {}
    for v in variants:
        err = call(v)
        if err.str is not None:
            raise RuntimeError(err.str.decode())
        samples = []
        for _ in range(reps):
            scrub.sum()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            call(v)
            e.record()
            torch.cuda.synchronize()
            samples.append(s.elapsed_time(e))
        times[v] = sorted(samples)[len(samples) // 2]
    return times


class Autotuner:
    def __init__(self, table=None):
        self.table = dict(table or {})

    @staticmethod
    def key(op, dtype, n_lists, n_elements):
        b = bucket(n_lists, n_elements)
        return f"{op}|{np.dtype(dtype).name}|{b[0]}|{b[1]}"

    def record(self, op, dtype, n_lists, n_elements, times):
        best = min(times, key=times.get)
        self.table[self.key(op, dtype, n_lists, n_elements)] = {"best": int(best),
                                                                 "ms": {str(k): v for k, v in times.items()}}
        return best

    def choose(self, op, n_lists, n_elements, itemsize, dtype=None):
        if dtype is not None:
            hit = self.table.get(self.key(op, dtype, n_lists, n_elements))
            if hit is not None:
                return hit["best"]
        return dispatch.heuristic_variant(n_lists, n_elements, itemsize)

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)

    @classmethod
    def load(cls, path):
        with open(path) as f:
            return cls(json.load(f))
