"""The hip backend's nplike: array creation and the array operations the
layout code performs outside the kernel registry, over torch tensors on the
current ROCm device.

It covers the subset of the reference's ``NumpyLike`` protocol
(``src/awkward/_nplikes/numpy_like.py:113-552``) that the jagged-array path
uses.  Buffers are torch tensors because PyTorch-ROCm owns device memory and
streams here; the data-parallel work of this path (gathers, boolean-mask
compaction, list lengths, elementwise ufuncs) runs in libawkward_hip.so
kernels; torch provides allocation and streams.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from awkward_amd._kernels import KernelError, device_address, set_stream
from awkward_amd._lib import lib
from awkward_amd.elementwise import DeviceExpr, build, compile_program, leaf

_TORCH = {
    np.dtype(np.bool_): torch.bool,
    np.dtype(np.int8): torch.int8,
    np.dtype(np.uint8): torch.uint8,
    np.dtype(np.int16): torch.int16,
    np.dtype(np.uint16): torch.uint16,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.uint32): torch.uint32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint64): torch.uint64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
}
_NUMPY = {v: k for k, v in _TORCH.items()}


def to_torch_dtype(dtype) -> torch.dtype:
    if isinstance(dtype, torch.dtype):
        return dtype
    return _TORCH[np.dtype(dtype)]


def to_numpy_dtype(dtype) -> np.dtype:
    if isinstance(dtype, torch.dtype):
        return _NUMPY[dtype]
    return np.dtype(dtype)


def _check(err: KernelError) -> None:
    if err.str is not None:
        raise RuntimeError(err.str.decode())


def _setup(name, nargs):
    f = getattr(lib(), name)
    f.argtypes = [ctypes.c_void_p] * (nargs - 1) + [ctypes.c_int64]
    f.restype = KernelError
    return f


class TorchHip:
    """NumpyLike over torch tensors on a ROCm device."""

    name = "hip"
    known_data = True

    def __init__(self, device: str | torch.device | None = None, fuse: bool = True):
        self.device = torch.device(device if device is not None else "cuda")
        # fuse=True: ufuncs return lazy DeviceExprs, evaluated in one pass
        # when needed (or inside the jagged mask getitem); False: each ufunc
        # writes its result at once, as the reference's nplike does
        self.fuse = fuse
        self._take = {}
        self._b2i = None
        self._sub = None
        self._eval = None
        self._mask = None

    # ---- creation --------------------------------------------------------
    def is_own_array(self, x) -> bool:
        return isinstance(x, torch.Tensor) and x.is_cuda

    def empty(self, n: int, dtype) -> torch.Tensor:
        return torch.empty(int(n), dtype=to_torch_dtype(dtype), device=self.device)

    def zeros(self, n: int, dtype) -> torch.Tensor:
        return torch.zeros(int(n), dtype=to_torch_dtype(dtype), device=self.device)

    def full(self, n: int, value, dtype) -> torch.Tensor:
        return torch.full((int(n),), value, dtype=to_torch_dtype(dtype), device=self.device)

    def arange(self, n: int, dtype=np.int64) -> torch.Tensor:
        return torch.arange(int(n), dtype=to_torch_dtype(dtype), device=self.device)

    def asarray(self, obj, dtype=None) -> torch.Tensor:
        if isinstance(obj, torch.Tensor):
            out = obj.to(self.device)
            if dtype is not None:
                out = out.to(to_torch_dtype(dtype))
            return out.contiguous()
        arr = np.ascontiguousarray(np.asarray(obj, dtype=None if dtype is None else to_numpy_dtype(dtype)))
        return torch.from_numpy(arr).to(self.device)

    def to_numpy(self, x) -> np.ndarray:
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
        return np.asarray(x)

    # ---- metadata ----------------------------------------------------------
    def dtype(self, x) -> np.dtype:
        if isinstance(x, DeviceExpr):
            return x.dtype
        return to_numpy_dtype(x.dtype)

    def length(self, x) -> int:
        if isinstance(x, DeviceExpr):
            return x.length
        return int(x.shape[0])

    def view(self, x, dtype):
        return x.view(to_torch_dtype(dtype))

    def item(self, x, i: int = 0):
        """Host read of one element (a synchronisation point)."""
        return x[i].item()

    def index_as_shape_item(self, x) -> int:
        return int(x.item())

    # ---- array operations ------------------------------------------------------
    def take(self, data: torch.Tensor, index: torch.Tensor) -> torch.Tensor:
        """data[index] for a 1-D contiguous data array and int64 index."""
        size = data.element_size()
        f = self._take.get(size)
        if f is None:
            f = self._take[size] = _setup(f"awkward_hip_take_{size}", 4)
        out = torch.empty(index.shape[0], dtype=data.dtype, device=self.device)
        set_stream()
        _check(f(device_address(out), device_address(data), device_address(index), index.shape[0]))
        return out

    def bool_to_jagged_index(self, offsets: torch.Tensor, mask: torch.Tensor):
        """(local indices of the true values, their offsets) for a jagged mask
        with int64 `offsets` over `mask`.  One synchronisation: the count."""
        if self._b2i is None:
            self._b2i = _setup("awkward_hip_bool_to_index", 5)
        n = offsets.shape[0] - 1
        toindex = torch.empty(max(int(mask.shape[0]), 1), dtype=torch.int64, device=self.device)
        tooffsets = torch.empty(n + 1, dtype=torch.int64, device=self.device)
        set_stream()
        _check(self._b2i(device_address(toindex), device_address(tooffsets), device_address(mask),
                         device_address(offsets), n))
        total = int(tooffsets[n].item())
        return toindex[:total], tooffsets

    def subtract(self, a, b):
        if a.dtype == torch.int64 and b.dtype == torch.int64 and a.is_contiguous() and b.is_contiguous():
            if self._sub is None:
                self._sub = _setup("awkward_hip_subtract_i64", 4)
            out = torch.empty_like(a)
            set_stream()
            _check(self._sub(device_address(out), device_address(a), device_address(b), a.shape[0]))
            return out
        return a - b

    def parents(self, offsets, total: int):
        """List index of every element under int64 `offsets`."""
        counts = (offsets[1:] - offsets[:-1]).to(torch.int64)
        return torch.repeat_interleave(torch.arange(counts.shape[0], device=self.device), counts,
                                       output_size=int(total))

    # ---- elementwise ------------------------------------------------------------
    def gathered(self, data, index):
        """data[index] as an elementwise operand: read in place by the
        expression (a lazy carry is never copied)."""
        if isinstance(data, DeviceExpr):
            data = data.materialize()
        return (data, index)

    def apply_ufunc(self, ufunc, *args):
        """A numpy ufunc over device columns (DeviceExprs, tensors, gathers)
        and scalars, with numpy's types: a lazy DeviceExpr when fusing, else
        written at once -- by the library's expression kernel either way."""
        expr = build(self, ufunc, args)
        return expr if self.fuse else expr.materialize()

    def eval_expr(self, expr: DeviceExpr) -> torch.Tensor:
        if self._eval is None:
            f = self._eval = getattr(lib(), "awkward_hip_expr_eval")
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
            f.restype = KernelError
        prog, keep = expr.program()
        out = torch.empty(max(expr.length, 0), dtype=to_torch_dtype(expr.dtype), device=self.device)
        set_stream()
        _check(self._eval(out.data_ptr() if expr.length else 0, ctypes.addressof(prog), 0, expr.length))
        del keep
        return out

    def mask_program(self, mask):
        """The program of a boolean mask: a recorded predicate, or a load of
        a boolean tensor."""
        if isinstance(mask, DeviceExpr):
            return mask.program()
        return compile_program(leaf(mask))

    def mask_getitem(self, fromstarts, fromstops, contentlen, moffsets, mask):
        """(tooffsets, tocarry) of array[mask] for a jagged boolean mask over
        int64 offsets `moffsets` and the indexed lists [fromstarts,
        fromstops) (awkward_hip_mask_getitem: the reference's mask -> index
        -> jagged_expand -> carrylen -> jagged_apply as the mask's bits and
        one scan).  The carry's length is tooffsets[n].  Raises the reference's
        IndexError message on a bad mask."""
        if self._mask is None:
            f = self._mask = getattr(lib(), "awkward_hip_mask_getitem")
            f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                                  ctypes.c_void_p]
            f.restype = KernelError
        n = moffsets.shape[0] - 1
        prog, keep = self.mask_program(mask)
        tooffsets = torch.empty(n + 1, dtype=torch.int64, device=self.device)
        tocarry = torch.empty(max(self.length(mask), 1), dtype=torch.int64, device=self.device)
        set_stream()
        err = self._mask(tooffsets.data_ptr(), tocarry.data_ptr(), fromstarts.data_ptr(), fromstops.data_ptr(),
                         int(contentlen), moffsets.data_ptr(), n, self.length(mask), ctypes.addressof(prog))
        del keep
        return err, tooffsets, tocarry
