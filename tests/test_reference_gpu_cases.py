"""The reference's own GPU-backend test cases for this path, run on the
oracle backend (CPU) and on the hip backend, each checked against a numpy
closed form of the same data.

* ``tests-cuda/test_3162_block_boundary_reducers.py``: 3,000 integers from
  ``default_rng(42).integers(3000)``, reduced flat and as three lists with
  offsets ``[0, 1, 2998, 3000]`` (a list crossing every block boundary)
  -- sum, sum of booleans, max, min, min of negated values, argmin, argmax,
  count.
* ``tests-cuda/test_cuda_nonlocal_reduce.py``: axis=0 of a ragged array,
  axis=1 of a three-deep array, int32 sums at axis=0 (int64 result, no
  overflow) and axis=0 with empty sublists -- sum and count.
"""

import numpy as np
import pytest

import awkward_amd as ak


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend()


BACKENDS = [pytest.param(oracle_backend, id="oracle"), pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu)]
OFFSETS = np.array([0, 1, 2998, 3000], dtype=np.int64)


def block_data(kind):
    rng = np.random.default_rng(seed=42)
    if kind == "bool":
        return rng.integers(2, size=3000, dtype=np.bool_)
    x = rng.integers(3000, size=3000)
    return -x if kind == "negative" else x


def closed_form(op, x, offsets):
    """numpy per list (None for an empty list under mask_identity)."""
    out = []
    for a, b in zip(offsets[:-1], offsets[1:]):
        v = x[a:b]
        if op == "count":
            out.append(int(b - a))
        elif op == "sum":
            out.append(int(v.astype(np.int64).sum()))
        elif b == a:
            out.append(None)
        else:
            out.append(int(getattr(np, op)(v)))
    return out


CASES = [("sum", "int"), ("sum", "bool"), ("max", "int"), ("min", "int"), ("min", "negative"), ("argmin", "int"),
         ("argmax", "int"), ("count", "int")]


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("op,kind", CASES)
def test_block_boundary(make, op, kind):
    be = make()
    x = block_data(kind)
    lists = ak.from_offsets(OFFSETS, x, be)
    flat = ak.Array(ak.from_offsets(np.array([0, len(x)]), x, be).layout.content)
    want_flat = closed_form(op, x, np.array([0, len(x)]))[0]
    assert getattr(ak, op)(flat, axis=-1) == want_flat
    assert getattr(ak, op)(lists, axis=-1).to_list() == closed_form(op, x, OFFSETS)


@pytest.mark.parametrize("make", BACKENDS)
def test_nonlocal_axis0_ragged(make):
    a = ak.from_list([[1.0, 2.0, 3.0], [4.0, 5.0], [6.0, 7.0, 8.0, 9.0]], make())
    assert ak.sum(a, axis=0).to_list() == [11.0, 14.0, 11.0, 9.0]
    assert ak.count(a, axis=0).to_list() == [3, 3, 2, 1]


@pytest.mark.parametrize("make", BACKENDS)
def test_nonlocal_three_deep_axis1(make):
    a = ak.from_list([[[1.0, 2], [3, 4, 5]], [[6.0], [7, 8]]], make())
    assert ak.sum(a, axis=1).to_list() == [[4.0, 6.0, 5.0], [13.0, 8.0]]
    assert ak.count(a, axis=1).to_list() == [[2, 2, 1], [2, 1]]


@pytest.mark.parametrize("make", BACKENDS)
def test_nonlocal_int32_axis0_no_overflow(make):
    a = ak.from_list([[100000, 200000], [300000, 400000, 500000]], make(), dtype=np.int32)
    out = ak.sum(a, axis=0)
    assert out.to_list() == [400000, 600000, 500000]
    assert np.dtype(out.layout.nplike.to_numpy(out.layout.data).dtype) == np.int64


@pytest.mark.parametrize("make", BACKENDS)
def test_nonlocal_axis0_empty_sublists(make):
    a = ak.from_list([[1.0, 2.0, 3.0], [], [6.0, 7.0]], make())
    assert ak.sum(a, axis=0).to_list() == [7.0, 9.0, 3.0]
    assert ak.count(a, axis=0).to_list() == [2, 2, 1]


@pytest.mark.parametrize("make", BACKENDS)
def test_axis_none_default(make):
    """axis=None is the reference's default: every number reduced to one
    value, missing values dropped, arg-reductions counting positions in the
    flattened numbers (``_do.py:225-340`` with ``completely_flatten``)."""
    be = make()
    rows = [[3, None, 7], [], None, [7, 1]]
    a = ak.from_list(rows, be, dtype=np.int64)
    flat = [3, 7, 7, 1]
    assert ak.sum(a) == 18 and ak.count(a) == 4
    assert ak.max(a) == 7 and ak.argmax(a) == flat.index(7)
    assert ak.min(a) == 1 and ak.argmin(a) == 3
    nested = ak.from_list([[[1.5, 2.5], []], [[4.0]]], be)
    assert ak.sum(nested) == 8.0 and ak.argmax(nested) == 2
    empty = ak.from_list([[], []], be, dtype=np.float64)
    assert ak.max(empty) is None and ak.sum(empty) == 0.0 and ak.count(empty) == 0


def test_completely_flatten_drops_missing_values():
    """The numbers axis=None reduces: every list level flattened in order,
    missing values (option types at any level) dropped."""
    from awkward_amd import contents as C

    be = oracle_backend()
    a = ak.from_list([[[1, None], []], None, [[2], [None, 3]]], be, dtype=np.int64)
    flat = C.completely_flatten(a.layout)
    assert flat.to_list() == [1, 2, 3]
    with pytest.raises(TypeError):
        C.completely_flatten(ak.from_offsets(np.array([0, 1]), {"x": np.array([1.0])}, be).layout)
