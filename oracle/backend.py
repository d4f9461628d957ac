"""A CPU backend over the oracle -- TEST INFRASTRUCTURE ONLY.

Runs the hip backend's own host layer (awkward_amd.highlevel / contents) with
numpy buffers and the oracle's kernels, so that whole pipelines (mask ->
jagged getitem -> range slice -> reductions, combinations, argsort) can be
checked end to end against the GPU, and timed as the CPU baseline.  The
nplike operations that are not registry kernels follow the reference's own
numpy steps (src/awkward/_slicing.py:476-530 for boolean masks,
numpyarray.py:367-373 for carries).
"""

from __future__ import annotations

import ctypes

import numpy as np

import oracle
from awkward_amd._kernel_signatures import SPECIALIZATIONS
from awkward_amd.backend import KernelBackend


class NumpyNplike:
    name = "cpu"
    known_data = True

    def is_own_array(self, x):
        return isinstance(x, np.ndarray)

    def empty(self, n, dtype):
        return np.empty(int(n), dtype=dtype)

    def zeros(self, n, dtype):
        return np.zeros(int(n), dtype=dtype)

    def full(self, n, value, dtype):
        return np.full(int(n), value, dtype=dtype)

    def arange(self, n, dtype=np.int64):
        return np.arange(int(n), dtype=dtype)

    def asarray(self, obj, dtype=None):
        if hasattr(obj, "detach"):
            obj = obj.detach().cpu().numpy()
        return np.ascontiguousarray(np.asarray(obj, dtype=dtype))

    def to_numpy(self, x):
        return np.asarray(x)

    def dtype(self, x):
        return x.dtype

    def length(self, x):
        return int(x.shape[0])

    def view(self, x, dtype):
        return x.view(dtype)

    def item(self, x, i=0):
        return x[i].item()

    def index_as_shape_item(self, x):
        return int(x[0])

    def take(self, data, index):
        return data[index]

    def bool_to_jagged_index(self, offsets, mask):
        # _slicing.py:476-530: local index, flatten, compact, cumsum, gather
        off = np.asarray(offsets, dtype=np.int64)
        n = len(off) - 1
        local = np.empty(off[-1] - off[0], dtype=np.int64)
        kernel = oracle.kernel("awkward_ListArray64_localindex_64")
        zero_based = off - off[0]
        err = kernel(local.ctypes.data, zero_based.ctypes.data, n)
        assert err.str is None
        flat_mask = mask[off[0]:off[-1]]
        nextcontent = local[flat_mask]
        cumsum = np.empty(len(flat_mask) + 1, dtype=np.int64)
        cumsum[0] = 0
        np.cumsum(flat_mask, out=cumsum[1:])
        return nextcontent, cumsum[zero_based]

    def subtract(self, a, b):
        return a - b

    def parents(self, offsets, total):
        return np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))

    def gathered(self, data, index):
        return data[index]

    def apply_ufunc(self, ufunc, *args):
        return ufunc(*args)


class OracleKernel:
    def __init__(self, name):
        self.name = name
        _, args, _ = SPECIALIZATIONS[name]
        self._pointer = ["List[" in t for _, t, _ in args]
        self._fn = oracle.kernel(name)

    def __call__(self, *args):
        call = []
        keep = []
        for a, p in zip(args, self._pointer):
            if not p:
                call.append(a)
            elif isinstance(a, np.ndarray):
                assert a.flags.c_contiguous
                call.append(a.ctypes.data)
            else:  # ctypes pointer table
                keep.append(a)
                call.append(ctypes.addressof(a))
        return self._fn(*call)


class OracleBackend(KernelBackend):
    name = "cpu-oracle"

    def __init__(self):
        super().__init__(NumpyNplike())

    def _make_kernel(self, name):
        return OracleKernel(name)

    def pointer_array(self, arrays):
        return (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
