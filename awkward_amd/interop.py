"""Zero-copy exchange of the hip backend's buffers: DLPack in, and the
Form + length + named-buffer triple of ``to_buffers`` / ``from_buffers``.

Every buffer of a hip layout is a torch tensor in the device's HBM, and a
torch tensor speaks DLPack (``__dlpack__`` / ``__dlpack_device__``), so what
leaves through ``to_buffers`` can be taken by any DLPack consumer without a
copy, and what arrives from a ROCm DLPack producer (another framework's
device array, an Arrow reader that landed columns in HBM) becomes a layout
buffer without a copy.

Reference behaviour followed:

* ``from_dlpack``: ``src/awkward/operations/ak_from_dlpack.py:14-90``
  (protocol check and its TypeError, device-type switch, ``regulararray``);
  the device codes are DLPack's ``DLDeviceType`` values, as listed in
  ``src/awkward/_connect/dlpack.py:9-19``.
* ``to_buffers``: ``src/awkward/_do.py:75-120`` (buffer keys
  ``"{form_key}-{attribute}"``, form keys ``"node{id}"`` numbered depth-first
  from ``id_start``, ``src/awkward/contents/content.py:206-235``) and each
  content's ``_to_buffers`` (e.g. ``contents/listoffsetarray.py:216-229``).
* ``from_buffers``: ``src/awkward/operations/ak_from_buffers.py:29`` and its
  ``_reconstitute`` / ``_from_buffer`` (the lengths each node implies for its
  buffers and content; a buffer shorter than the form needs is a TypeError;
  an own array is viewed and truncated, not copied).
* Form dictionaries: ``src/awkward/forms/*form.py`` ``_to_dict_part``.

This package has one backend, ``hip``: the reference sends a CPU or pinned
host buffer to its CPU backend, here it is copied once into device memory
(``prefer_cpu`` is accepted for the same call signature and has no CPU
backend to select).  Offsets, starts, stops and index buffers of 32-bit
type are widened to int64 on the device, because the jagged kernels of this
backend are the 64-bit specializations.
"""

from __future__ import annotations

import builtins
import json
from enum import IntEnum

import numpy as np
import torch

from awkward_amd import contents as C
from awkward_amd.nplike import to_numpy_dtype, to_torch_dtype

__all__ = ("DLPackDevice", "from_dlpack", "to_buffers", "from_buffers")


class DLPackDevice(IntEnum):
    """DLPack's ``DLDeviceType`` codes (the subset the reference names)."""

    CPU = 1
    CUDA = 2
    CUDA_PINNED = 3
    OPENCL = 4
    VULKAN = 7
    METAL = 8
    VPI = 9
    ROCM = 10
    ROCM_PINNED = 11
    CUDA_MANAGED = 13


_DEVICE_RESIDENT = (DLPackDevice.ROCM, DLPackDevice.CUDA)
_HOST_READABLE = (DLPackDevice.CPU, DLPackDevice.CUDA_PINNED, DLPackDevice.ROCM_PINNED, DLPackDevice.CUDA_MANAGED)

_INDEX_DTYPES = {"i8": np.int8, "u8": np.uint8, "i32": np.int32, "u32": np.uint32, "i64": np.int64}
_INDEX_NAMES = {np.dtype(v): k for k, v in _INDEX_DTYPES.items()}


def _backend(backend):
    if backend is None or backend == "hip":
        from awkward_amd.backend import hip_backend

        return hip_backend()
    if isinstance(backend, str):
        raise ValueError(f"unknown backend {backend!r}; this package provides 'hip'")
    return backend


def _wrap(layout, highlevel):
    if not highlevel:
        return layout
    from awkward_amd.highlevel import Array

    return Array(layout)


def _check_dtype(dtype):
    try:
        to_torch_dtype(to_numpy_dtype(dtype))
    except (KeyError, TypeError) as err:
        raise TypeError(f"dtype {dtype} is not supported by the hip backend") from err


# ---- DLPack in ----------------------------------------------------------------
def _tensor_from_dlpack(array, nplike):
    """The DLPack producer's memory as a torch tensor on `nplike`'s device:
    shared when the producer is on that ROCm device, copied once from host
    memory otherwise."""
    try:
        dlpack_device = array.__dlpack_device__
    except AttributeError as err:
        raise TypeError(
            f"Expected an object that implements the DLPack protocol, received {type(array)}"
        ) from err
    device_type, device_id = dlpack_device()
    try:
        device_type = DLPackDevice(int(device_type))
    except ValueError as err:
        raise ValueError(f"DLPack device type {device_type} is not supported") from err

    if device_type in _DEVICE_RESIDENT:
        tensor = torch.from_dlpack(array)
        target = nplike.device
        if target.index is not None and tensor.device.index != target.index:
            raise ValueError(
                f"DLPack buffer is on device {device_id}, the hip backend on {target}; "
                "move it first (cross-device buffers are not shared implicitly)"
            )
        return tensor
    if device_type in _HOST_READABLE:
        return torch.from_dlpack(array).to(nplike.device)
    raise ValueError(f"DLPack device type {device_type.name} is not supported")


def _regular_from(tensor, backend):
    """Nested RegularArrays over a flat NumpyArray, one per inner dimension
    (``from_arraylib`` with ``regulararray=True``)."""
    shape = [int(s) for s in tensor.shape]
    layout = C.NumpyArray(tensor.contiguous().reshape(-1), backend)
    for k in range(len(shape) - 1, 0, -1):
        layout = C.RegularArray(layout, shape[k], int(np.prod(shape[:k])), backend)
    return layout


def from_dlpack(array, *, prefer_cpu=True, regulararray=False, highlevel=True, backend=None):
    """An Array over the DLPack producer's buffer
    (``src/awkward/operations/ak_from_dlpack.py:14-90``).

    A ROCm (or CUDA-typed, as some producers report HIP memory) device
    buffer on the backend's device is shared, not copied.  CPU, pinned and
    managed buffers are copied into device memory, since this package has no
    CPU backend; `prefer_cpu` is accepted for the reference's signature.
    A 0-d array is rejected, as the reference's default
    ``primitive_policy="error"`` does.  Multidimensional arrays become a
    NumpyArray with inner dimensions, or nested RegularArrays when
    `regulararray` is True.
    """
    del prefer_cpu
    if not hasattr(array, "__dlpack_device__"):
        raise TypeError(f"Expected an object that implements the DLPack protocol, received {type(array)}")
    backend = _backend(backend)
    tensor = _tensor_from_dlpack(array, backend.nplike)
    _check_dtype(tensor.dtype)
    if tensor.dim() == 0:
        raise TypeError("cannot convert a 0-dimensional (scalar) array into an Array")
    if tensor.dim() > 1 and regulararray:
        return _wrap(_regular_from(tensor, backend), highlevel)
    return _wrap(C.NumpyArray(tensor, backend), highlevel)


# ---- to_buffers ---------------------------------------------------------------
def _index_name(tensor):
    dtype = to_numpy_dtype(tensor.dtype)
    if dtype == np.bool_:
        return "i8"
    try:
        return _INDEX_NAMES[dtype]
    except KeyError as err:
        raise TypeError(f"index buffer of dtype {dtype} has no Form index type") from err


def _index_buffer(tensor):
    return tensor.view(torch.int8) if tensor.dtype == torch.bool else tensor


def to_buffers(
    array, container=None, buffer_key="{form_key}-{attribute}", form_key="node{id}", id_start=0, backend="cpu"
):
    """``(form, length, container)`` for a hip layout
    (``src/awkward/operations/ak_to_buffers.py``, ``src/awkward/_do.py:75-120``).

    With ``backend="cpu"`` (the reference's default) the container holds
    NumPy copies in main memory.  With ``backend=None`` or ``"hip"`` it holds
    the layout's own device tensors, not copies, each exportable through
    DLPack.  An IndexedArray is written as it stands (its index and its
    whole content), as the reference does."""
    layout = getattr(array, "layout", array)
    if backend not in (None, "hip", "cpu"):
        raise ValueError(f"unknown backend {backend!r}; expected 'cpu', 'hip' or None")
    to_host = backend == "cpu"
    if container is None:
        container = {}
    if form_key is None:
        raise TypeError("a 'form_key' must be supplied, to match Form elements to buffers in the 'container'")
    counter = [int(id_start)]

    def getkey(key, attribute):
        if isinstance(buffer_key, str):
            return buffer_key.format(form_key=key, attribute=attribute)
        if callable(buffer_key):
            return buffer_key(form_key=key, attribute=attribute)
        raise TypeError(f"buffer_key must be a string or a callable, not {type(buffer_key)}")

    def put(name, tensor):
        container[name] = tensor.detach().cpu().numpy() if to_host else tensor

    def nextkey():
        key = form_key.format(id=counter[0]) if isinstance(form_key, str) else form_key(id=counter[0])
        counter[0] += 1
        return key

    def walk(node):
        key = nextkey()
        if isinstance(node, C.NumpyArray):
            data = node.data
            form = {"class": "NumpyArray", "primitive": str(to_numpy_dtype(data.dtype))}
            if data.dim() > 1:
                form["inner_shape"] = [int(s) for s in data.shape[1:]]
            form["form_key"] = key
            put(getkey(key, "data"), data.contiguous().reshape(-1))
            return form
        if isinstance(node, C.ListOffsetArray):
            form = {"class": "ListOffsetArray", "offsets": _index_name(node.offsets)}
            put(getkey(key, "offsets"), node.offsets)
            form["content"] = walk(node.content)
        elif isinstance(node, C.ListArray):
            form = {"class": "ListArray", "starts": _index_name(node.starts), "stops": _index_name(node.stops)}
            put(getkey(key, "starts"), node.starts)
            put(getkey(key, "stops"), node.stops)
            form["content"] = walk(node.content)
        elif isinstance(node, C.RegularArray):
            form = {"class": "RegularArray", "size": int(node.size), "content": walk(node.content)}
        elif isinstance(node, C.RecordArray):
            form = {"class": "RecordArray", "fields": None if node.fields is None else list(node.fields)}
            form["contents"] = [walk(c) for c in node.contents]
        elif isinstance(node, C.IndexedOptionArray):
            form = {"class": "IndexedOptionArray", "index": _index_name(node.index)}
            put(getkey(key, "index"), node.index)
            form["content"] = walk(node.content)
        elif isinstance(node, C.IndexedArray):
            form = {"class": "IndexedArray", "index": _index_name(node.index)}
            put(getkey(key, "index"), node.index)
            form["content"] = walk(node.content)
        elif isinstance(node, C.ByteMaskedArray):
            form = {"class": "ByteMaskedArray", "mask": _index_name(node.mask), "valid_when": bool(node.valid_when)}
            put(getkey(key, "mask"), _index_buffer(node.mask))
            form["content"] = walk(node.content)
        else:
            raise NotImplementedError(f"to_buffers of {type(node).__name__}")
        form["form_key"] = key
        return form

    form = walk(layout)
    return form, layout.length, container


# ---- from_buffers -------------------------------------------------------------
def _as_tensor(raw, dtype, count, key, nplike):
    """``count`` items of `dtype` from one container value
    (``ak_from_buffers.py`` ``_from_buffer``): a device tensor or DLPack
    producer is viewed as `dtype` and truncated; bytes or a host array are
    reinterpreted and copied to the device."""
    if callable(raw) and not hasattr(raw, "__dlpack__"):
        raw = raw()
    tdtype = to_torch_dtype(dtype)
    if isinstance(raw, torch.Tensor) or hasattr(raw, "__dlpack_device__"):
        if not isinstance(raw, torch.Tensor):
            raw = _tensor_from_dlpack(raw, nplike)
        elif raw.device != nplike.device and not (raw.is_cuda and nplike.device.type == "cuda"):
            raw = raw.to(nplike.device)
        flat = raw.contiguous().reshape(-1)
        if flat.dtype != tdtype:
            try:
                flat = flat.view(tdtype)
            except RuntimeError as err:
                raise TypeError(f"buffer {key!r} of dtype {flat.dtype} cannot be viewed as {np.dtype(dtype)}") from err
        if flat.shape[0] < count:
            raise TypeError(f"size of array ({flat.shape[0]}) is less than size of form ({count})")
        return flat[:count]
    host = np.frombuffer(memoryview(raw).cast("B"), dtype=dtype)
    if host.shape[0] < count:
        raise TypeError(f"size of array ({host.shape[0]}) is less than size of form ({count})")
    return nplike.asarray(np.array(host[:count]))  # a writable host copy for the upload


def _to_int64(x):
    return x if x.dtype == torch.int64 else x.to(torch.int64)


def _reconstitute(form, length, container, getkey, backend):
    nplike = backend.nplike
    if isinstance(form, str):
        form = {"class": "NumpyArray", "primitive": form}
    cls = form["class"]
    key = form.get("form_key")
    params = form.get("parameters") or {}
    if "__array__" in params:
        raise NotImplementedError(f"{cls} with __array__={params['__array__']!r} is not on the hip backend's path")

    def buffer(attribute, dtype, count):
        name = getkey(key, attribute)
        if name not in container:
            raise KeyError(f"container has no buffer {name!r}")
        return _as_tensor(container[name], dtype, count, name, nplike)

    if cls == "NumpyArray":
        dtype = np.dtype(form["primitive"])
        _check_dtype(dtype)
        inner = [int(s) for s in form.get("inner_shape", [])]
        width = int(np.prod(inner)) if inner else 1
        data = buffer("data", dtype, length * width)
        if inner:
            data = data.reshape(length, *inner)
        return C.NumpyArray(data, backend)
    if cls == "ListOffsetArray":
        offsets = buffer("offsets", _INDEX_DTYPES[form["offsets"]], length + 1)
        offsets = _to_int64(offsets)
        inner = int(offsets[-1]) if length >= 0 and offsets.shape[0] > 0 else 0
        return C.ListOffsetArray(offsets, _reconstitute(form["content"], inner, container, getkey, backend), backend)
    if cls == "ListArray":
        starts = _to_int64(buffer("starts", _INDEX_DTYPES[form["starts"]], length))
        stops = _to_int64(buffer("stops", _INDEX_DTYPES[form["stops"]], length))
        inner = 0 if length == 0 else builtins.max(0, int(stops.max()))
        return C.ListArray(starts, stops, _reconstitute(form["content"], inner, container, getkey, backend), backend)
    if cls == "RegularArray":
        size = int(form["size"])
        content = _reconstitute(form["content"], length * size, container, getkey, backend)
        return C.RegularArray(content, size, length, backend)
    if cls == "RecordArray":
        fields = form.get("fields")
        contents = [_reconstitute(c, length, container, getkey, backend) for c in form["contents"]]
        return C.RecordArray(contents, fields, length, backend)
    if cls == "IndexedArray":
        index = _to_int64(buffer("index", _INDEX_DTYPES[form["index"]], length))
        inner = 0 if length == 0 else builtins.max(0, int(index.max()) + 1)
        return C.IndexedArray(index, _reconstitute(form["content"], inner, container, getkey, backend), backend)
    if cls == "IndexedOptionArray":
        index = buffer("index", _INDEX_DTYPES[form["index"]], length)
        inner = 0 if length == 0 else builtins.max(0, int(index.max()) + 1)
        return C.IndexedOptionArray(index, _reconstitute(form["content"], inner, container, getkey, backend), backend)
    if cls == "ByteMaskedArray":
        mask = buffer("mask", _INDEX_DTYPES[form["mask"]], length)
        content = _reconstitute(form["content"], length, container, getkey, backend)
        return C.ByteMaskedArray(mask, content, bool(form["valid_when"]), backend)
    raise NotImplementedError(f"from_buffers of a {cls} node")


def from_buffers(form, length, container, buffer_key="{form_key}-{attribute}", *, backend="hip", highlevel=True):
    """An Array from a Form, a length and named buffers
    (``src/awkward/operations/ak_from_buffers.py:29``), the inverse of
    ``to_buffers``.  `form` is a Form dictionary or its JSON text; container
    values may be device tensors or DLPack producers (shared), host arrays
    or bytes (copied), or zero-argument callables returning one of those
    (called once, here)."""
    backend = _backend(backend)
    if isinstance(form, str) and form.lstrip().startswith("{"):
        form = json.loads(form)
    length = int(length)
    if length < 0:
        raise ValueError("length must be non-negative")

    def getkey(key, attribute):
        if isinstance(buffer_key, str):
            return buffer_key.format(form_key=key, attribute=attribute)
        return buffer_key(form_key=key, attribute=attribute)

    return _wrap(_reconstitute(form, length, container, getkey, backend), highlevel)
