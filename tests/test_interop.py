"""DLPack ingest and the to_buffers / from_buffers triple on the hip backend.

Pinned against the reference's own DLPack tests
(``tests/test_2649_dlpack_support.py``: memory shared with the producer,
the ``DLPackOf`` wrapper, the TypeError for a non-DLPack argument) and the
worked example in ``src/awkward/operations/ak_to_buffers.py``'s docstring
(the exact Form dictionary, length and buffer keys for
``[[1, 2, 3], [], [4, 5]]``).
"""

import numpy as np
import pytest

import awkward_amd as ak
from awkward_amd.interop import DLPackDevice


class DLPackOf:
    """Forwards only the DLPack protocol of the wrapped array."""

    def __init__(self, array):
        self._array = array

    def __dlpack__(self, stream=None):
        return self._array.__dlpack__() if stream is None else self._array.__dlpack__(stream=stream)

    def __dlpack_device__(self):
        return self._array.__dlpack_device__()


# ---- CPU: argument checks and the device codes --------------------------------
def test_invalid_argument():
    with pytest.raises(TypeError, match=r"Expected an object that implements the DLPack protocol"):
        ak.from_dlpack([1, 2, 3])


def test_device_codes_match_producers():
    """DLPack's DLDeviceType values, as reported by real producers."""
    import torch

    assert np.arange(3).__dlpack_device__()[0] == DLPackDevice.CPU
    assert torch.arange(3).__dlpack_device__()[0] == DLPackDevice.CPU
    assert (DLPackDevice.CUDA, DLPackDevice.ROCM, DLPackDevice.ROCM_PINNED) == (2, 10, 11)


def test_to_buffers_rejects_unknown_backend():
    with pytest.raises(ValueError, match="unknown backend"):
        ak.to_buffers(None, backend="jax")


# ---- GPU ----------------------------------------------------------------------
def _hip():
    from awkward_amd.backend import hip_backend

    return hip_backend()


@pytest.mark.gpu
def test_from_dlpack_device_tensor_shares_memory():
    import torch

    t = torch.arange(1000, dtype=torch.float64, device="cuda")
    arr = ak.from_dlpack(t)
    assert arr.layout.data.data_ptr() == t.data_ptr()
    t[7] = -1.0
    assert arr.to_list()[7] == -1.0
    arr2 = ak.from_dlpack(DLPackOf(t))
    assert arr2.layout.data.data_ptr() == t.data_ptr()


@pytest.mark.gpu
def test_from_dlpack_numpy_is_copied_to_device():
    np_array = np.arange(2 * 3 * 4 * 5).reshape(2, 3, 4, 5)
    arr = ak.from_dlpack(np_array)
    assert arr.layout.data.is_cuda
    assert arr.to_list() == np_array.tolist()
    reg = ak.from_dlpack(DLPackOf(np_array), regulararray=True)
    from awkward_amd import contents as C

    node, sizes = reg.layout, []
    while isinstance(node, C.RegularArray):
        sizes.append((node.size, node.length))
        node = node.content
    assert sizes == [(3, 2), (4, 6), (5, 24)]
    assert reg.to_list() == np_array.tolist()


@pytest.mark.gpu
def test_from_dlpack_rejects_scalars_and_unsupported_dtypes():
    import torch

    with pytest.raises(TypeError, match="0-dimensional"):
        ak.from_dlpack(torch.tensor(3.0, device="cuda"))
    with pytest.raises(TypeError, match="not supported"):
        ak.from_dlpack(torch.zeros(4, dtype=torch.complex64, device="cuda"))


@pytest.mark.gpu
def test_to_buffers_docstring_example():
    """The reference docstring's worked example, to the key and the value."""
    original = ak.from_list([[1, 2, 3], [], [4, 5]], _hip(), dtype=np.int64)
    form, length, container = ak.to_buffers(original)
    assert form == {
        "class": "ListOffsetArray",
        "offsets": "i64",
        "content": {"class": "NumpyArray", "primitive": "int64", "form_key": "node1"},
        "form_key": "node0",
    }
    assert length == 3
    assert sorted(container) == ["node0-offsets", "node1-data"]
    np.testing.assert_array_equal(container["node0-offsets"], [0, 3, 3, 5])
    np.testing.assert_array_equal(container["node1-data"], [1, 2, 3, 4, 5])
    back = ak.from_buffers(form, length, container)
    assert back.to_list() == [[1, 2, 3], [], [4, 5]]


@pytest.mark.gpu
def test_to_buffers_device_container_is_zero_copy():
    import json

    off = np.array([0, 2, 2, 5, 9], dtype=np.int64)
    rng = np.random.default_rng(3)
    content = {"pt": rng.random(9).astype(np.float32), "q": rng.integers(-1, 2, 9).astype(np.int8)}
    arr = ak.from_offsets(off, content, _hip())
    form, length, container = ak.to_buffers(arr, backend=None, id_start=5, form_key="n{id}", buffer_key="{form_key}/{attribute}")
    assert form["form_key"] == "n5" and form["content"]["form_key"] == "n6"
    assert [c["form_key"] for c in form["content"]["contents"]] == ["n7", "n8"]
    assert sorted(container) == ["n5/offsets", "n7/data", "n8/data"]
    assert container["n5/offsets"].data_ptr() == arr.layout.offsets.data_ptr()
    back = ak.from_buffers(json.dumps(form), length, container, buffer_key="{form_key}/{attribute}")
    assert back.layout.offsets.data_ptr() == arr.layout.offsets.data_ptr()
    assert back.to_list() == arr.to_list()
    assert ak.sum(back["pt"], axis=-1).to_list() == ak.sum(arr["pt"], axis=-1).to_list()


@pytest.mark.gpu
def test_round_trip_every_node_type():
    import torch

    from awkward_amd import contents as C

    be = _hip()
    dev = be.nplike.device
    leaf = C.NumpyArray(torch.arange(12, dtype=torch.float64, device=dev), be)
    layouts = [
        C.ListArray(torch.tensor([0, 4, 4], device=dev), torch.tensor([3, 4, 12], device=dev), leaf, be),
        C.RegularArray(leaf, 4, 3, be),
        C.IndexedArray(torch.tensor([11, 0, 5], device=dev), leaf, be),
        C.ByteMaskedArray(torch.tensor([1, 0, 1], dtype=torch.int8, device=dev), C.NumpyArray(leaf.data[:3], be), True, be),
        C.RecordArray([leaf, C.NumpyArray(leaf.data * 2, be)], None, 12, be),
    ]
    for layout in layouts:
        for backend in ("cpu", None):
            form, length, container = ak.to_buffers(layout, backend=backend)
            back = ak.from_buffers(form, length, container)
            assert back.to_list() == layout.to_list(), (type(layout).__name__, backend)


@pytest.mark.gpu
def test_from_buffers_host_bytes_and_narrow_offsets():
    """Raw bytes and 32-bit offsets: the offsets are widened to int64 on the
    device, and the jagged reductions run on the result."""
    from oracle.backend import OracleBackend

    rng = np.random.default_rng(11)
    counts = rng.poisson(5, 300)
    off = np.zeros(301, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    data = rng.normal(size=int(off[-1]))
    form = {
        "class": "ListOffsetArray",
        "offsets": "i32",
        "content": {"class": "NumpyArray", "primitive": "float64", "form_key": "node1"},
        "form_key": "node0",
    }
    container = {"node0-offsets": off.astype(np.int32).tobytes(), "node1-data": data.tobytes()}
    hip = ak.from_buffers(form, 300, container)
    assert hip.layout.offsets.dtype.itemsize == 8
    ref = ak.from_offsets(off, data, OracleBackend())
    for fn in (ak.max, ak.argmax, ak.count):
        assert fn(hip, axis=-1).to_list() == fn(ref, axis=-1).to_list(), fn.__name__
    assert ak.num(hip).to_list() == ak.num(ref).to_list()


@pytest.mark.gpu
def test_from_buffers_errors():
    form = {
        "class": "ListOffsetArray",
        "offsets": "i64",
        "content": {"class": "NumpyArray", "primitive": "float64", "form_key": "node1"},
        "form_key": "node0",
    }
    short = {"node0-offsets": np.array([0, 1, 2], dtype=np.int64), "node1-data": np.zeros(1)}
    with pytest.raises(TypeError, match="less than size of form"):
        ak.from_buffers(form, 2, short)
    with pytest.raises(KeyError, match="node1-data"):
        ak.from_buffers(form, 2, {"node0-offsets": np.array([0, 1, 2], dtype=np.int64)})
