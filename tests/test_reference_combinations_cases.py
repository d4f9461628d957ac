"""The reference's GPU-backend combinations cases
(``tests-cuda/test_3260_combinations_n_equal_3.py``,
``tests-cuda/test_3150_combinations_n_equal_2.py``): the ragged array
``[[0.0, 1.1, 2.2, 3.3], [], [4.4, 5.5, 6.6], [7.7], [8.8 ... 12.2]]`` and a
2 x 4 RegularArray, n = 2 and 3, with and without replacement, with fields,
and argcombinations -- on the oracle backend and on the hip backend, against
the reference's literals and ``itertools``."""

import itertools

import numpy as np
import pytest

import awkward_amd as ak
from awkward_amd import contents as C

ROWS = [[0.0, 1.1, 2.2, 3.3], [], [4.4, 5.5, 6.6], [7.7], [8.8, 9.9, 10.0, 11.1, 12.2]]


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend()


BACKENDS = [pytest.param(oracle_backend, id="oracle"), pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu)]


def expected(rows, n, replacement):
    f = itertools.combinations_with_replacement if replacement else itertools.combinations
    return [list(f(r, n)) for r in rows]


@pytest.mark.parametrize("make", BACKENDS)
def test_literals_n3(make):
    a = ak.from_list(ROWS, make())
    got = ak.combinations(a, 3, replacement=False).to_list()
    assert got[:4] == [[(0.0, 1.1, 2.2), (0.0, 1.1, 3.3), (0.0, 2.2, 3.3), (1.1, 2.2, 3.3)], [], [(4.4, 5.5, 6.6)], []]
    assert ak.argcombinations(a, 3, replacement=False).to_list()[:4] == [
        [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)], [], [(0, 1, 2)], []]


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("replacement", [False, True])
def test_ragged_against_itertools(make, n, replacement):
    a = ak.from_list(ROWS, make())
    assert ak.combinations(a, n, replacement=replacement).to_list() == expected(ROWS, n, replacement)
    names = ["x", "y", "z"][:n]
    with_fields = ak.combinations(a, n, replacement=replacement, fields=names).to_list()
    assert with_fields == [[dict(zip(names, t)) for t in row] for row in expected(ROWS, n, replacement)]
    idx = [list(range(len(r))) for r in ROWS]
    assert ak.argcombinations(a, n, replacement=replacement).to_list() == expected(idx, n, replacement)


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("n", [2, 3])
def test_regular_array(make, n):
    be = make()
    flat = np.array([0.0, 1.1, 2.2, 3.3, 4.4, 5.5, 6.6, 7.7])
    r = ak.Array(C.RegularArray(C.NumpyArray(be.nplike.asarray(flat), be), 4, 2, be))
    rows = [list(flat[:4]), list(flat[4:])]
    for replacement in (False, True):
        assert ak.combinations(r, n, replacement=replacement).to_list() == expected(rows, n, replacement)
