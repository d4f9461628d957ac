"""The reference's reducer tests for result dtypes and boolean reductions,
run on the oracle backend (CPU) and on the hip backend against the
reference's literals -- not one backend against the other:

* ``tests-cuda/test_3136_cuda_reducers.py:28-260`` and
  ``tests/test_0115_generic_reducer_operation.py:1408-1600``
  (``sumprod_types``): the result dtype of ``ak.sum`` / ``ak.prod`` at
  axis=-1 equals numpy's ``np.sum`` / ``np.prod`` of the same 2-D array
  (numpy evaluated here, so the promotion rules of the installed numpy pin
  the host layer's), and the totals agree;
* ``test_0115``'s ``test_any`` (:1603), ``test_all`` (:1639), the boolean
  blocks of ``test_count_min`` (:1782-1816) and ``test_count_max``
  (:1852-1884) at axes -1, 1, -2 and 0, and ``test_prod_bool`` (:2243-2270);
* the depth-3 axis handling of ``ak.num`` / ``ak.local_index`` /
  ``ak.combinations`` / ``ak.argsort`` at axis=-1 (``ak_num.py``,
  ``ak_local_index.py``, ``ak_combinations.py``: the axis counts from the
  numbers, as ``regularize_axis`` makes it).
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


def plain_hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend(fused=False)


BACKENDS = [
    pytest.param(oracle_backend, id="oracle"),
    pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu),
    pytest.param(plain_hip_backend, id="hip-registry", marks=pytest.mark.gpu),
]
OFFSETS = np.array([0, 3, 3, 5, 6], dtype=np.int64)


def result_dtype(out):
    layout = out.layout
    while not hasattr(layout, "data"):  # through a ByteMaskedArray of mask_identity=True
        layout = layout.content
    return np.asarray(layout.nplike.to_numpy(layout.data)).dtype


def product(xs):
    out = 1
    for x in xs:
        out *= x
    return out


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("dtype", [np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
                                   np.uint64])
def test_sumprod_types(make, dtype):
    if dtype is np.bool_:
        array = np.array([[True, False, False], [True, False, False]])
    else:
        array = np.array([[0, 1, 2], [3, 4, 5]], dtype=dtype)
    depth1 = ak.from_offsets(OFFSETS, array.reshape(-1), make())
    s = ak.sum(depth1, axis=-1)
    p = ak.prod(depth1, axis=-1)
    assert result_dtype(s) == np.sum(array, axis=-1).dtype
    assert result_dtype(p) == np.prod(array, axis=-1).dtype
    assert sum(np.sum(array, axis=-1).tolist()) == sum(s.to_list())
    assert product(np.prod(array, axis=-1).tolist()) == product(p.to_list())


ANY_X = np.array([1.1, 2.2, 3.3, 0.0, 2.2, 0.0, 0.0, 0.0, 0.0, 0.0])
ALL_X = np.array([1.1, 2.2, 3.3, 0.0, 2.2, 0.0, 0.0, 2.2, 0.0, 4.4])
MIN_B = np.array([True, True, True, False, True, False, False, True, False, True])
MAX_B = np.array([False, True, True, False, True, False, False, False, False, False])
OFF3 = np.array([0, 3, 6, 10], dtype=np.int64)


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("op,x,inner,outer", [
    ("any", ANY_X, [True, True, False], [True, True, True, False]),
    ("all", ALL_X, [True, False, False], [False, True, False, True]),
    ("min", MIN_B, [True, False, False], [False, True, False, True]),
    ("max", MAX_B, [True, True, False], [False, True, True, False]),
])
def test_bool_reducers_every_axis(make, op, x, inner, outer):
    depth1 = ak.from_offsets(OFF3, x, make())
    fn = getattr(ak, op)
    for axis, want in ((-1, inner), (1, inner), (-2, outer), (0, outer)):
        out = fn(depth1, axis)
        assert out.to_list() == want, (op, axis)
        assert result_dtype(out) == np.bool_


@pytest.mark.parametrize("make", BACKENDS)
def test_prod_bool(make):
    array = np.array([[True, False, False], [True, False, False]])
    depth1 = ak.from_offsets(OFFSETS, array.reshape(-1), make())
    assert ak.prod(depth1, axis=-1).to_list() == [0, 1, 0, 0]
    assert ak.all(depth1, axis=-1).to_list() == [False, True, False, False]
    assert ak.min(depth1, axis=-1).to_list() == [False, None, False, False]
    depth1 = ak.from_offsets(OFFSETS, array.view(np.uint8).reshape(-1), make())
    assert ak.prod(depth1, axis=-1).to_list() == [0, 1, 0, 0]
    assert ak.all(depth1, axis=-1).to_list() == [0, 1, 0, 0]
    assert ak.min(depth1, axis=-1).to_list() == [0, None, 0, 0]


@pytest.mark.parametrize("make", BACKENDS)
def test_bool_max_min_mask_and_empty(make):
    """max/min of booleans keep mask_identity=True (None for an empty list)
    and give the all-false / all-true identity without it."""
    a = ak.from_list([[True, False], [], [False], [True, True]], make())
    assert ak.max(a, axis=-1).to_list() == [True, None, False, True]
    assert ak.min(a, axis=-1).to_list() == [False, None, False, True]
    assert ak.max(a, axis=-1, mask_identity=False).to_list() == [True, False, False, True]
    assert ak.min(a, axis=-1, mask_identity=False).to_list() == [False, True, False, True]
    assert ak.any(a, axis=-1).to_list() == [True, False, False, True]
    assert ak.all(a, axis=-1).to_list() == [False, True, False, True]


@pytest.mark.parametrize("make", BACKENDS)
def test_any_all_nan_and_signed_zero(make):
    """x != 0: NaN is nonzero and -0.0 is zero (the CPU kernels' test)."""
    a = ak.from_list([[np.nan], [-0.0], [0.0, np.nan], [-0.0, 1.0]], make())
    assert ak.any(a, axis=-1).to_list() == [True, False, True, True]
    assert ak.all(a, axis=-1).to_list() == [True, False, False, False]


@pytest.mark.parametrize("make", BACKENDS)
def test_axis_minus_one_depth3(make):
    a = ak.from_list([[[1, 2], [3]], [[4]]], make(), dtype=np.int64)
    assert ak.num(a, axis=-1).to_list() == [[2, 1], [1]]
    assert ak.num(a, axis=1).to_list() == [2, 1]
    assert ak.local_index(a, axis=-1).to_list() == [[[0, 1], [0]], [[0]]]
    assert ak.local_index(a, axis=1).to_list() == [[0, 1], [0]]
    assert ak.combinations(a, 2, axis=-1).to_list() == [[[(1, 2)], []], [[]]]
    assert ak.argsort(a, axis=-1, ascending=False).to_list() == [[[1, 0], [0]], [[0]]]
    assert ak.sort(a, axis=2, ascending=False).to_list() == [[[2, 1], [3]], [[4]]]
    with pytest.raises(ValueError):
        ak.num(a, axis=3)
    with pytest.raises(NotImplementedError):
        ak.argsort(a, axis=1)
