"""Pins of the oracle beyond the reference's golden vectors: closed forms and
library routines that a plausible mistake (a dropped term, a wrong sign or
index, a swapped tie-break) would contradict."""

import itertools
import math

import numpy as np
import pytest

from parity import offsets_from_counts

import oracle


def call(name, *args):
    conv = []
    keep = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(a.ctypes.data)
        elif isinstance(a, list):
            import ctypes

            t = (ctypes.c_void_p * len(a))(*[x.ctypes.data for x in a])
            keep.append(t)
            conv.append(ctypes.addressof(t))
        else:
            conv.append(a)
    return oracle.kernel(name)(*conv)


@pytest.fixture
def rng():
    return np.random.default_rng(99)


def test_sum_int_closed_form(rng):
    counts = rng.poisson(7, 2000)
    off = offsets_from_counts(counts)
    x = rng.integers(-(2**62), 2**62, int(off[-1]), dtype=np.int64)
    out = np.zeros(len(counts), np.int64)
    call("awkward_reduce_sum_int64_int64_64", out, x, off, len(counts))
    want = [int(sum(int(v) for v in x[a:b])) for a, b in zip(off[:-1], off[1:])]
    want = [((w + 2**63) % 2**64) - 2**63 for w in want]  # two's-complement wrap
    assert out.tolist() == want


def test_sum_float_fsum_bound(rng):
    counts = rng.poisson(50, 500)
    off = offsets_from_counts(counts)
    x = rng.standard_normal(int(off[-1]))
    out = np.zeros(len(counts))
    call("awkward_reduce_sum_float64_float64_64", out, x, off, len(counts))
    for i, (a, b) in enumerate(zip(off[:-1], off[1:])):
        exact = math.fsum(x[a:b])
        assert abs(out[i] - exact) <= (b - a) * np.finfo(np.float64).eps * np.abs(x[a:b]).sum() + 1e-300
        # serial left-to-right order, not pairwise
        acc = 0.0
        for v in x[a:b]:
            acc += v
        assert out[i] == acc


def test_sum_float32_accumulates_in_float32(rng):
    x = np.full(20000, 0.1, dtype=np.float32)
    off = np.array([0, 20000], dtype=np.int64)
    out = np.zeros(1, np.float32)
    call("awkward_reduce_sum_float32_float32_64", out, x, off, 1)
    acc = np.float32(0)
    for v in x:
        acc = np.float32(acc + v)
    assert out[0] == acc and out[0] != np.float32(2000.0)


def test_argmax_matches_numpy_without_nan(rng):
    counts = rng.poisson(9, 3000)
    off = offsets_from_counts(counts)
    x = rng.integers(0, 5, int(off[-1])).astype(np.float64)  # many ties
    out = np.zeros(len(counts), np.int64)
    call("awkward_reduce_argmax_float64_64", out, x, off, off[:-1].copy(), len(counts))
    want = [a + int(np.argmax(x[a:b])) if b > a else -1 for a, b in zip(off[:-1], off[1:])]
    assert out.tolist() == want
    call("awkward_reduce_argmin_float64_64", out, x, off, off[:-1].copy(), len(counts))
    want = [a + int(np.argmin(x[a:b])) if b > a else -1 for a, b in zip(off[:-1], off[1:])]
    assert out.tolist() == want


def test_extremum_special_values():
    x = np.array([np.nan, -0.0, 0.0, np.nan, 0.0, -0.0, -np.inf], dtype=np.float64)
    off = np.array([0, 1, 3, 4, 6, 7], dtype=np.int64)
    out = np.zeros(5)
    call("awkward_reduce_max_float64_float64_64", out, x, off, 5, -np.inf)
    # NaN never selected; the first of -0.0/+0.0 kept; empty -> identity
    assert np.isneginf(out[0]) and np.signbit(out[1]) and np.isneginf(out[2])
    assert not np.signbit(out[3]) and np.isneginf(out[4])
    arg = np.zeros(5, np.int64)
    call("awkward_reduce_argmax_float64_64", arg, x, off, off[:-1].copy(), 5)
    assert arg.tolist() == [0, 1, 3, 4, 6]  # a NaN first element wins


@pytest.mark.parametrize("n,replacement", [(1, False), (2, False), (3, False), (2, True), (3, True)])
def test_combinations_itertools(rng, n, replacement):
    counts = rng.integers(0, 9, 300)
    off = offsets_from_counts(counts)
    starts, stops = off[:-1].copy(), off[1:].copy()
    total = np.zeros(1, np.int64)
    tooff = np.zeros(len(counts) + 1, np.int64)
    call("awkward_ListArray64_combinations_length_64", total, tooff, n, replacement, starts, stops, len(counts))
    gen = itertools.combinations_with_replacement if replacement else itertools.combinations
    want = [t for a, b in zip(starts, stops) for t in gen(range(a, b), n)]
    assert int(total[0]) == len(want)
    sizes = [math.comb(int(c) + (n - 1 if replacement else 0), n) for c in counts]
    assert tooff.tolist() == [0] + np.cumsum(sizes).tolist()
    carry = [np.zeros(len(want), np.int64) for _ in range(n)]
    toindex = np.zeros(n, np.int64)
    call("awkward_ListArray64_combinations_64", carry, toindex, np.zeros(n, np.int64), n, replacement, starts, stops,
         len(counts))
    got = list(zip(*[c.tolist() for c in carry]))
    assert got == want and toindex.tolist() == [len(want)] * n


NONE = 2**63 - 1


@pytest.mark.parametrize("sl", [(1, None, 1), (None, None, -1), (-3, None, 2), (5, 1, -2), (None, -2, 3), (-1, -8, -3),
                                (2, 2, 1), (10, None, 1)])
def test_next_range_python_slices(rng, sl):
    counts = rng.integers(0, 12, 400)
    off = offsets_from_counts(counts)
    fs, fe = off[:-1].copy(), off[1:].copy()
    start, stop, step = sl
    a = NONE if start is None else start
    b = NONE if stop is None else stop
    cl = np.zeros(1, np.int64)
    call("awkward_ListArray64_getitem_next_range_carrylength", cl, fs, fe, len(counts), a, b, step)
    want = [s0 + j for s0, c in zip(fs, counts) for j in range(*slice(start, stop, step).indices(int(c)))]
    assert int(cl[0]) == len(want)
    tooff = np.zeros(len(counts) + 1, np.int64)
    carry = np.zeros(max(len(want), 1), np.int64)
    call("awkward_ListArray64_getitem_next_range_64", tooff, carry, fs, fe, len(counts), a, b, step)
    assert carry[: len(want)].tolist() == want


def nan_first_key(x, ascending):
    k = np.where(np.isnan(x), -np.inf, x if ascending else -x)
    return np.where(np.isnan(x), 0, 1), k


@pytest.mark.parametrize("ascending", [True, False])
def test_argsort_lexsort(rng, ascending):
    counts = rng.integers(0, 30, 300)
    off = offsets_from_counts(counts)
    x = rng.integers(-3, 4, int(off[-1])).astype(np.float64)
    x[rng.random(len(x)) < 0.1] = np.nan
    x[rng.random(len(x)) < 0.1] = -0.0
    out = np.zeros(len(x), np.int64)
    call("awkward_argsort_float64_int64", out, x, len(x), off, len(off), ascending, True)
    for a, b in zip(off[:-1], off[1:]):
        seg = x[a:b]
        isnum, k = nan_first_key(seg, ascending)
        k = np.where(k == 0, 0.0, k)  # -0.0 ties +0.0
        want = np.lexsort((np.arange(b - a), k, isnum))
        assert out[a:b].tolist() == want.tolist()


def test_argsort_reference_literal():
    # tests/test_0945_argsort_sort_nan_array.py:12-25
    x = np.array([1, 2, np.nan, 3, 0, np.nan], dtype=np.float64)
    out = np.zeros(6, np.int64)
    off = np.array([0, 6], dtype=np.int64)
    call("awkward_argsort_float64_int64", out, x, 6, off, 2, True, True)
    assert out.tolist() == [2, 5, 4, 0, 1, 3]


def test_sort_descending_is_not_reversed_ascending():
    x = np.array([1, 1, 0, 1], dtype=np.int64)
    off = np.array([0, 4], dtype=np.int64)
    out = np.zeros(4, np.int64)
    call("awkward_argsort_int64_int64", out, x, 4, off, 2, False, True)
    assert out.tolist() == [0, 1, 3, 2]  # stable: ties keep input order


def test_jagged_apply_first_error_in_loop_order():
    # list 1 has an out-of-range index, list 2 a bad slice; list 1 is reported
    slicestarts = np.array([0, 1, 2], np.int64)
    slicestops = np.array([1, 2, 1], np.int64)
    sliceindex = np.array([0, 7, 0], np.int64)
    fs = np.array([0, 2, 4], np.int64)
    fe = np.array([2, 4, 6], np.int64)
    err = call("awkward_ListArray64_getitem_jagged_apply_64", np.zeros(4, np.int64), np.zeros(3, np.int64),
               slicestarts, slicestops, 3, sliceindex, 3, fs, fe, 6)
    assert err.str == b"index out of range" and err.id == 1 and err.attempt == 7


def test_planted_mistake_is_caught(rng):
    """A plausible bug planted in a copy of the restatement (argmax ties to
    the LAST occurrence) is rejected by the same checks."""
    counts = rng.poisson(9, 500)
    off = offsets_from_counts(counts)
    x = rng.integers(0, 3, int(off[-1])).astype(np.float64)

    def wrong_argmax(x, off):
        out = []
        for a, b in zip(off[:-1], off[1:]):
            best = -1
            for i in range(a, b):
                if best == -1 or x[i] >= x[best]:
                    best = i
            out.append(best)
        return out

    want = [a + int(np.argmax(x[a:b])) if b > a else -1 for a, b in zip(off[:-1], off[1:])]
    assert wrong_argmax(x, off) != want
