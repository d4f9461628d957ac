"""The ``hip`` backend: the kernel registry and error handling.

Mirrors the reference's ``Backend`` ABC (``src/awkward/_backends/backend.py:24-62``):
``backend[(name, *dtypes)]`` returns a callable kernel, and
``maybe_kernel_error`` raises ``ValueError`` with the message format of
``Backend.format_kernel_error`` (``backend.py:38-56``).  Keys are built with the
reference generator's rule: the kernel family name followed by the dtype of
every pointer argument in specification order
(``dev/generate-kernel-signatures.py:225-236``).
"""

from __future__ import annotations

import numpy as np

from awkward_amd._kernel_signatures import SPECIALIZATIONS

kSliceNone = 2**63 - 1  # awkward-cpp/include/awkward/common.h:51

_BY_KEY = {key: name for name, (_, _, key) in SPECIALIZATIONS.items()}


def _dtype_name(x) -> str:
    if isinstance(x, str):
        return np.dtype(x).name if x not in ("float", "double") else {"float": "float32", "double": "float64"}[x]
    try:
        import torch

        if isinstance(x, torch.dtype):
            from awkward_amd.nplike import to_numpy_dtype

            return to_numpy_dtype(x).name
    except ImportError:  # pragma: no cover
        pass
    return np.dtype(x).name


def normalize_key(key) -> tuple:
    name, *dtypes = key
    return (name, *(_dtype_name(d) for d in dtypes))


def specialization_for(key) -> str:
    k = normalize_key(key)
    try:
        return _BY_KEY[k]
    except KeyError:
        raise KeyError(f"no hip kernel for {k}") from None


class KernelBackend:
    """Shared registry logic; subclasses supply the kernel class and nplike."""

    name = "abstract"

    def __init__(self, nplike):
        self.nplike = nplike
        self.index_nplike = nplike
        self._cache = {}

    def _make_kernel(self, name: str):
        raise NotImplementedError

    def __getitem__(self, key):
        k = normalize_key(key)
        kernel = self._cache.get(k)
        if kernel is None:
            kernel = self._cache[k] = self._make_kernel(specialization_for(k))
        return kernel

    def prepare_reducer(self, reducer):
        return reducer

    def format_kernel_error(self, error) -> str:
        filename = "" if error.filename is None else " (in compiled code: " + error.filename.decode(
            errors="surrogateescape").lstrip("\n").lstrip("(")
        message = error.str.decode(errors="surrogateescape")
        if error.attempt != kSliceNone:
            message += f" while attempting to get index {error.attempt}"
        return message + filename

    def maybe_kernel_error(self, error) -> None:
        if error is None or error.str is None:
            return
        raise ValueError(self.format_kernel_error(error))


class HipBackend(KernelBackend):
    """``fused=True`` (default): the layout code calls one tuned kernel per
    reduction (see awkward_amd.reducers); ``fused=False`` issues the
    reference's exact kernel sequence through the registry."""

    name = "hip"

    def __init__(self, device=None, fused=True):
        from awkward_amd.nplike import TorchHip

        super().__init__(TorchHip(device, fuse=fused))
        self.fused = fused

    def _make_kernel(self, name: str):
        from awkward_amd._kernels import HipKernel

        return HipKernel(name)

    def pointer_array(self, arrays):
        """Device array of the device addresses of `arrays` (combinations' tocarry)."""
        from awkward_amd._kernels import PointerArray

        return PointerArray(arrays)


_instance = None


def hip_backend() -> HipBackend:
    global _instance
    if _instance is None:
        _instance = HipBackend()
    return _instance
