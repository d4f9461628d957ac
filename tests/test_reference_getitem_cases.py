"""The reference's jagged-getitem literals, on the oracle backend (CPU) and
on the hip backend (GPU), each against the literal -- not one backend
against the other:

* ``tests/test_0111_jagged_and_masked_getitem.py``: ``test_jagged``
  (:668-683), ``test_double_jagged`` (:686-698, the array-by-array part),
  ``test_jagged_mask`` (:1186-1244), ``test_array_boolean_to_int``
  (:1346-1380, the cases without None: the boolean leaves as local
  positions), ``test_record`` (:845-864, the slice without None),
  ``test_numpyarray`` (:835-842) and ``test_emptyarray`` (:808-832, the
  slices without None);
* the "cannot fit jagged slice" length check (listarray.py:494-501).

Option-typed slices (None inside the slice) and RegularArray heads are not
on the hip path (they raise NotImplementedError)."""

import numpy as np
import pytest

import awkward_amd as ak


def backends():
    from oracle.backend import OracleBackend

    yield pytest.param("oracle", id="oracle")
    yield pytest.param("hip", id="hip", marks=pytest.mark.gpu)


def make(name):
    if name == "oracle":
        from oracle.backend import OracleBackend

        return OracleBackend()
    from awkward_amd.backend import HipBackend

    return HipBackend()


def A(be, x, dtype=None):
    return ak.from_list(x, be, dtype)


FLOATS = [[1.1, 2.2, 3.3], [], [4.4, 5.5], [6.6], [7.7, 8.8, 9.9]]


@pytest.mark.parametrize("name", backends())
def test_jagged(name):
    be = make(name)
    array = A(be, FLOATS)
    array2 = A(be, [[0, -1], [], [-1, 0], [-1], [1, 1, -2, 0]])
    assert array[array2].to_list() == [[1.1, 3.3], [], [5.5, 4.4], [6.6], [8.8, 8.8, 8.8, 7.7]]


@pytest.mark.parametrize("name", backends())
def test_double_jagged(name):
    be = make(name)
    array = A(be, [[[0, 1, 2, 3], [4, 5]], [[6, 7, 8], [9, 10, 11, 12, 13]]])
    array2 = A(be, [[[2, 1, 0], [-1]], [[-1, -2, -3], [2, 1, 1, 3]]])
    assert array[array2].to_list() == [[[2, 1, 0], [5]], [[8, 7, 6], [11, 10, 10, 12]]]
    # an inner list of the slice longer than the array's: the descend check
    with pytest.raises(IndexError):
        array[A(be, [[[2, 1, 0], [-1], [0]], [[-1, -2, -3], [2, 1, 1, 3]]])]


MASKS = [
    ([[True, True, True], [], [True, True], [True], [True, True, True]], FLOATS),
    ([[False, True, True], [], [True, True], [True], [True, True, True]],
     [[2.2, 3.3], [], [4.4, 5.5], [6.6], [7.7, 8.8, 9.9]]),
    ([[True, False, True], [], [True, True], [True], [True, True, True]],
     [[1.1, 3.3], [], [4.4, 5.5], [6.6], [7.7, 8.8, 9.9]]),
    ([[True, True, True], [], [False, True], [True], [True, True, True]],
     [[1.1, 2.2, 3.3], [], [5.5], [6.6], [7.7, 8.8, 9.9]]),
    ([[True, True, True], [], [False, False], [True], [True, True, True]],
     [[1.1, 2.2, 3.3], [], [], [6.6], [7.7, 8.8, 9.9]]),
]


@pytest.mark.parametrize("name", backends())
@pytest.mark.parametrize("case", range(len(MASKS)))
def test_jagged_mask(name, case):
    be = make(name)
    mask, want = MASKS[case]
    assert A(be, FLOATS)[A(be, mask)].to_list() == want


BOOL_TO_INT = [
    ([[True, True, True], [], [True, True], [True], [True, True, True, True]],
     [[0, 1, 2], [], [0, 1], [0], [0, 1, 2, 3]]),
    ([[True, True, False], [], [True, False], [False], [True, True, True, False]], [[0, 1], [], [0], [], [0, 1, 2]]),
    ([[False, True, True], [], [False, True], [False], [False, True, True, True]], [[1, 2], [], [1], [], [1, 2, 3]]),
]


@pytest.mark.parametrize("name", backends())
@pytest.mark.parametrize("case", range(len(BOOL_TO_INT)))
def test_array_boolean_to_int(name, case):
    """The boolean leaves as the local positions of their true values
    (_slicing._normalise_item_bool_to_int), and the getitem they drive."""
    be = make(name)
    mask, want = BOOL_TO_INT[case]
    m = A(be, mask).layout
    nplike = be.nplike
    index, offsets = nplike.bool_to_jagged_index(m.offsets, m.content.data)
    idx, off = nplike.to_numpy(index), nplike.to_numpy(offsets)
    assert [idx[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)] == want
    values = [[10 * i + j for j in range(len(row))] for i, row in enumerate(mask)]
    got = A(be, values, np.int64)[A(be, mask)].to_list()
    assert got == [[values[i][j] for j in w] for i, w in enumerate(want)]


@pytest.mark.parametrize("name", backends())
def test_record(name):
    be = make(name)
    array = A(be, [{"x": [0, 1, 2], "y": [0.0, 1.1, 2.2, 3.3]}, {"x": [3, 4, 5, 6], "y": [4.4, 5.5]},
                   {"x": [7, 8], "y": [6.6, 7.7, 8.8, 9.9]}])
    array2 = A(be, [[-1, 1], [0, 0, 1], [-1, -2]])
    got = array[array2].layout
    assert got.contents[0].to_list() == [[2, 1], [3, 3, 4], [8, 7]]
    assert got.contents[1].to_list() == [[3.3, 1.1], [4.4, 4.4, 5.5], [9.9, 8.8]]


@pytest.mark.parametrize("name", backends())
def test_numpyarray_too_deep(name):
    be = make(name)
    array = A(be, [[1.1, 2.2, 3.3], [], [4.4, 5.5]])
    with pytest.raises(IndexError):
        array[A(be, [[[], [], []], [], [[], []]])]


@pytest.mark.parametrize("name", backends())
def test_emptyarray(name):
    be = make(name)
    empty = A(be, [[], [], [], []])
    assert empty[A(be, [[], [], [], []])].to_list() == [[], [], [], []]
    with pytest.raises(IndexError):
        empty[A(be, [[], [0], [], []])]


@pytest.mark.parametrize("name", backends())
def test_cannot_fit_jagged_slice(name):
    be = make(name)
    with pytest.raises(IndexError, match="cannot fit jagged slice"):
        A(be, FLOATS)[A(be, [[0], [], [0]])]
    with pytest.raises(IndexError, match="cannot fit jagged slice"):
        A(be, FLOATS)[A(be, [[True], [], [True]])]
