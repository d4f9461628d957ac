"""Option-type contents under the reducers: lists of optional numbers
(IndexedOptionArray / ByteMaskedArray inside the lists) and missing lists
(an option type over the lists).

Pinned by the reference's own literals (``tests/test_0992_correct_ptp_unmasking.py``
sum/min at axis=-1; ``tests/test_0070_argmin_and_argmax.py::test_jagged``,
IndexedOptionArray over lists; ``tests/test_0115_generic_reducer_operation.py::
test_ByteMaskedArray`` mask layout), by closed forms in plain Python (the
reductions of the non-None values, argmin/argmax positions counting the
Nones), and, on the GPU, by the oracle on the same seeded inputs -- every
new kernel bit-exact.
"""

import numpy as np
import pytest

import awkward_amd as ak
from awkward_amd import contents as C
from parity import offsets_from_counts, run_both, same_bits


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend()


BACKENDS = [pytest.param(oracle_backend, id="oracle"), pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu)]

PTP_ROWS = [[0, 1, None], [None, 3], [], None, [4, 5, None, 6]]


@pytest.mark.parametrize("make", BACKENDS)
def test_reference_literals_missing_values_and_lists(make):
    be = make()
    a = ak.from_list(PTP_ROWS, be)
    assert ak.sum(a, axis=-1).to_list() == [1, 3, 0, None, 15]
    assert ak.min(a, axis=-1).to_list() == [0, 3, None, None, 4]
    assert ak.max(a, axis=-1).to_list() == [1, 3, None, None, 6]
    assert ak.argmin(a, axis=-1).to_list() == [0, 1, None, None, 0]
    assert ak.argmax(a, axis=-1).to_list() == [1, 1, None, None, 3]
    assert ak.count(a, axis=-1).to_list() == [2, 1, 0, None, 3]


@pytest.mark.parametrize("make", BACKENDS)
def test_reference_literal_option_over_lists(make):
    be = make()
    nplike = be.nplike
    v2 = ak.from_list([[2.2, 1.1, 3.3], [], [4.4, 5.5], [5.5], [-4.4, -5.5, -6.6]], be).layout
    io = C.IndexedOptionArray(nplike.asarray(np.array([4, 3, -1, 4, 0], dtype=np.int64)), v2, be)
    assert ak.argmin(ak.Array(io), axis=-1).to_list() == [2, 0, None, 2, 1]
    bm = C.ByteMaskedArray(nplike.asarray(np.array([0, 0, 1, 1, 0], dtype=np.int8)), v2, False, be)
    assert ak.argmin(ak.Array(bm), axis=-1).to_list() == [1, None, None, None, 2]


def closed_forms(rows):
    def arg(row, better):
        best = None
        for i, v in enumerate(row):
            if v is not None and (best is None or better(v, row[best])):
                best = i
        return best

    vals = [[v for v in r if v is not None] for r in rows]
    return {
        "sum": [sum(v) for v in vals],
        "max": [max(v) if v else None for v in vals],
        "min": [min(v) if v else None for v in vals],
        "argmax": [arg(r, lambda a, b: a > b) for r in rows],
        "argmin": [arg(r, lambda a, b: a < b) for r in rows],
        "count": [len(v) for v in vals],
    }


def random_rows(seed, n=400):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        k = int(rng.poisson(5))
        row = [int(x) for x in rng.integers(-50, 50, k)]
        for j in range(k):
            if rng.random() < 0.3:
                row[j] = None
        rows.append(row)
    return rows


def as_bytemasked(rows, be):
    """The same lists with a ByteMaskedArray(valid_when=True) inside."""
    flat = [v for r in rows for v in r]
    mask = np.array([v is not None for v in flat], dtype=np.int8)
    data = np.array([0 if v is None else v for v in flat], dtype=np.int64)
    off = offsets_from_counts([len(r) for r in rows])
    nplike = be.nplike
    inner = C.ByteMaskedArray(nplike.asarray(mask), C.NumpyArray(nplike.asarray(data), be), True, be)
    return ak.Array(C.ListOffsetArray(nplike.asarray(off), inner, be))


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("kind", ["indexed", "bytemasked"])
def test_reducers_against_closed_forms(make, kind):
    be = make()
    rows = random_rows(7)
    a = ak.from_list(rows, be, dtype=np.int64) if kind == "indexed" else as_bytemasked(rows, be)
    want = closed_forms(rows)
    for name in ("sum", "max", "min", "argmax", "argmin", "count"):
        assert getattr(ak, name)(a, axis=-1).to_list() == want[name], name


# ---- kernel level, GPU vs oracle at sizes beyond one scan tile -----------------
def _mask(rng, n, p=0.3):
    return (rng.random(n) >= p).astype(np.int8) * rng.integers(1, 5, n).astype(np.int8)


@pytest.mark.gpu
@pytest.mark.parametrize("validwhen", [True, False])
def test_bytemasked_kernels(validwhen):
    rng = np.random.default_rng(31)
    counts = rng.poisson(6, 40000)
    counts[::997] = rng.integers(3000, 9000, len(counts[::997]))
    off = offsets_from_counts(counts)
    n = int(off[-1])
    mask = _mask(rng, n)
    valid = int(((mask != 0) == validwhen).sum())
    h, d, he, de = run_both("awkward_ByteMaskedArray_numnull", [np.zeros(1, np.int64), mask, n, validwhen])
    assert he.str is None and de.str is None and same_bits(h[0], d[0]) and int(h[0][0]) == n - valid
    nl = len(counts)
    h, d, he, de = run_both("awkward_ByteMaskedArray_reduce_next_64",
                            [np.zeros(valid, np.int64), np.zeros(nl + 1, np.int64), np.zeros(n, np.int64), mask, off,
                             nl, validwhen])
    assert he.str is None and de.str is None
    for k in range(3):
        assert same_bits(h[k], d[k]), k
    shifts = rng.integers(0, 100, n).astype(np.int64)
    for name, extra in (("awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64", []),
                        ("awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64", [shifts])):
        h, d, he, de = run_both(name, [np.zeros(valid, np.int64), mask, n, validwhen] + extra)
        assert he.str is None and de.str is None and same_bits(h[0], d[0]), name
    h, d, _, _ = run_both("awkward_ByteMaskedArray_getitem_nextcarry_64", [np.zeros(valid, np.int64), mask, n, validwhen])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both("awkward_ByteMaskedArray_getitem_nextcarry_outindex_64",
                          [np.zeros(valid, np.int64), np.zeros(n, np.int64), mask, n, validwhen])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])


@pytest.mark.gpu
@pytest.mark.parametrize("P,ctype", [("32", np.int32), ("U32", np.uint32), ("64", np.int64)])
def test_indexed_kernels(P, ctype):
    rng = np.random.default_rng(32)
    counts = rng.poisson(6, 30000)
    off = offsets_from_counts(counts)
    n = int(off[-1])
    lencontent = 5000
    index = rng.integers(0, lencontent, n).astype(np.int64)
    if ctype != np.uint32:
        index[rng.random(n) < 0.3] = -1
    index = index.astype(ctype)
    valid = int((index.astype(np.int64) >= 0).sum())
    h, d, he, de = run_both(f"awkward_IndexedArray{P}_numnull", [np.zeros(1, np.int64), index, n])
    assert he.str is None and same_bits(h[0], d[0])
    nl = len(counts)
    h, d, he, de = run_both(f"awkward_IndexedArray{P}_reduce_next_64",
                            [np.zeros(valid, np.int64), np.zeros(nl + 1, np.int64), np.zeros(n, np.int64), index, off, nl])
    assert he.str is None and de.str is None
    for k in range(3):
        assert same_bits(h[k], d[k]), k
    shifts = rng.integers(0, 100, n).astype(np.int64)
    h, d, _, _ = run_both(f"awkward_IndexedArray{P}_reduce_next_nonlocal_nextshifts_64", [np.zeros(valid, np.int64), index, n])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both(f"awkward_IndexedArray{P}_reduce_next_nonlocal_nextshifts_fromshifts_64",
                          [np.zeros(valid, np.int64), index, n, shifts])
    assert same_bits(h[0], d[0])
    h, d, he, de = run_both(f"awkward_IndexedArray{P}_getitem_nextcarry_outindex_64",
                            [np.zeros(valid, np.int64), np.zeros(n, ctype), index, n, lencontent])
    assert he.str is None and de.str is None and same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    # an index past the content: the first one in loop order is the error
    bad = index.copy()
    bad[[1234, 20000]] = ctype(lencontent + 3)
    h, d, he, de = run_both(f"awkward_IndexedArray{P}_getitem_nextcarry_outindex_64",
                            [np.zeros(valid, np.int64), np.zeros(n, ctype), bad, n, lencontent])
    assert he.str is not None and (he.str, he.id, he.attempt) == (de.str, de.id, de.attempt)


@pytest.mark.gpu
def test_adjust_starts_shifts():
    rng = np.random.default_rng(33)
    n = 50000
    toptr = rng.integers(-1, 20000, n).astype(np.int64)
    starts = rng.integers(0, 100, n).astype(np.int64)
    shifts = rng.integers(0, 100, 20000).astype(np.int64)
    h, d, _, _ = run_both("awkward_NumpyArray_reduce_adjust_starts_shifts_64",
                          [toptr, n, np.zeros(n + 1, np.int64), starts, shifts])
    assert same_bits(h[0], d[0])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["indexed", "bytemasked"])
def test_reducers_hip_vs_oracle_large(kind):
    """Host layer on both backends over 20k lists with 30% missing values."""
    rows = random_rows(8, n=20000)
    results = {}
    for label, be in (("oracle", oracle_backend()), ("hip", hip_backend())):
        a = ak.from_list(rows, be, dtype=np.int64) if kind == "indexed" else as_bytemasked(rows, be)
        results[label] = {name: getattr(ak, name)(a, axis=-1).to_list()
                          for name in ("sum", "max", "min", "argmax", "argmin", "count")}
    assert results["hip"] == results["oracle"]
