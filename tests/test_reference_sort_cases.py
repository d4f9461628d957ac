"""The reference's GPU-backend sort test cases
(``tests-cuda/test_4240_cuda_argsort.py``: every dtype on a ragged array
with empty lists, duplicates, NaN, ±inf and signed zeros; stable argsort
indices, gathering by the argsort, sort; NaN before a real -inf; all-empty
lists; 32-bit offsets), run on the oracle backend and on the hip backend and
checked against a closed form of the CPU kernels' order
(``awkward_argsort.cpp:23-60``: NaN first in both directions, -0.0 == +0.0,
ties in input order)."""

import math

import numpy as np
import pytest

import awkward_amd as ak

INTEGER_DTYPES = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64"]
FLOAT_DTYPES = ["float32", "float64"]
NAN, INF = float("nan"), float("inf")


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend()


BACKENDS = [pytest.param(oracle_backend, id="oracle"), pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu)]


def rows_for(dtype):
    if dtype == "bool":
        return [[True, False, True, True], [], [False, False], [True]]
    if dtype in FLOAT_DTYPES:
        return [[3.0, NAN, 1.0, 1.0], [], [NAN, NAN, 4.0], [2.0], [9.0, 0.0, 0.0, 7.0, 7.0], [-INF, NAN, 0.0, INF],
                [INF, -INF, -0.0, 0.0, NAN, -2.5]]
    return [[3, 1, 2, 1], [], [5, 5, 4], [2], [9, 0, 0, 7, 7]]


def closed_form_argsort(row, ascending):
    """NaN first (in input order), then the values, ties in input order."""
    nans = [i for i, v in enumerate(row) if isinstance(v, float) and math.isnan(v)]
    rest = [i for i, v in enumerate(row) if not (isinstance(v, float) and math.isnan(v))]
    rest.sort(key=lambda i: (row[i] if ascending else -float(row[i])))  # list.sort is stable
    return nans + rest


def same(a, b):
    if isinstance(a, list):
        return isinstance(b, list) and len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    if isinstance(a, float) and isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or a == b
    return a == b


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("dtype", ["bool", *INTEGER_DTYPES, *FLOAT_DTYPES])
@pytest.mark.parametrize("ascending", [True, False])
def test_argsort_sort_carry(make, dtype, ascending):
    rows = rows_for(dtype)
    a = ak.from_list(rows, make(), dtype=np.dtype(dtype))
    host = [[np.dtype(dtype).type(v).item() for v in r] for r in rows]
    want = [closed_form_argsort(r, ascending) for r in host]
    got = ak.argsort(a, axis=-1, ascending=ascending, stable=True)
    assert got.to_list() == want
    want_sorted = [[r[i] for i in idx] for r, idx in zip(host, want)]
    assert same(a[got].to_list(), want_sorted)  # gathering by the argsort
    assert same(ak.sort(a, axis=-1, ascending=ascending, stable=True).to_list(), want_sorted)


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("ascending", [True, False])
def test_nan_before_real_infinities(make, ascending):
    a = ak.from_list([[-INF, NAN, 0.0, INF]], make())
    want = [[1, 0, 2, 3]] if ascending else [[1, 3, 2, 0]]
    assert ak.argsort(a, axis=-1, ascending=ascending, stable=True).to_list() == want


@pytest.mark.parametrize("make", BACKENDS)
def test_all_empty(make):
    a = ak.from_list([[], [], []], make(), dtype=np.float64)
    assert ak.argsort(a, axis=-1).to_list() == [[], [], []]


@pytest.mark.gpu
def test_int32_offsets_kernel():
    """awkward_argsort_float64_int32 with 32-bit offsets: [3, 1, 2] -> [1, 2, 0], [5, 4] -> [1, 0]."""
    from parity import run_both, same_bits

    x = np.array([3.0, 1.0, 2.0, 5.0, 4.0])
    off = np.array([0, 3, 5], dtype=np.int32)
    h, d, he, de = run_both("awkward_argsort_float64_int32", [np.zeros(5, np.int64), x, 5, off, 3, True, True])
    assert he.str is None and de.str is None
    assert h[0].tolist() == [1, 2, 0, 1, 0] and same_bits(h[0], d[0])
