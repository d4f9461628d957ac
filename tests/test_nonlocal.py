"""Reductions above the innermost lists (axis != -1): the non-local branch.

Pinned by the reference's own literals:
``tests/test_0410_fix_argminmax_positions_for_missing_values.py``
(``test_jagged_axis0``, ``test_jagged_axis1``: argmin positions counting the
empty sublists before them), ``tests/test_0070_argmin_and_argmax.py``
(``test_2d``, ``test_3d`` at every axis), ``tests/test_0115_generic_reducer_operation.py``
(min/max of the primes layout at axis=-1/-2) and
``tests/test_0992_correct_ptp_unmasking.py`` (sum/min at axis=-2 with
missing values and lists); by a pure-Python closed form on random
three-level arrays; and, on the GPU, by the oracle on the same inputs with
every non-local kernel bit-exact.
"""

import numpy as np
import pytest

import awkward_amd as ak
from parity import offsets_from_counts, run_both, same_bits


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def hip_backend():
    from awkward_amd.backend import HipBackend

    return HipBackend()


BACKENDS = [pytest.param(oracle_backend, id="oracle"), pytest.param(hip_backend, id="hip", marks=pytest.mark.gpu)]

AXIS1_CASES = [
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]]], [[4, 3, 2], [4, 3, 2]]),
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[], [1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]]], [[4, 3, 2], [5, 4, 3]]),
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[], [], [1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]]], [[4, 3, 2], [6, 5, 4]]),
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[1.1], [1.1, 2.2], [], [1.1, 2.2, 3.3], [999, 2.0], [1.0]]], [[4, 3, 2], [5, 4, 3]]),
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [], [999, 2.0], [1.0]]], [[4, 3, 2], [5, 4, 2]]),
    ([[[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [1.0]],
      [[1.1], [1.1, 2.2], [1.1, 2.2, 3.3], [999, 2.0], [], [1.0]]], [[4, 3, 2], [5, 3, 2]]),
]

ARRAY_3D = [
    [[3.3, 2.2, 5.5, 1.1, 4.4], [4.4, 2.2, 1.1, 3.3, 5.5], [2.2, 1.1, 4.4, 3.3, 5.5]],
    [[-3.3, 2.2, -5.5, 1.1, 4.4], [4.4, -2.2, 1.1, -3.3, 5.5], [2.2, 1.1, 4.4, 3.3, -5.5]],
]


@pytest.mark.parametrize("make", BACKENDS)
def test_jagged_axis0(make):
    a = ak.from_list([[1.1, 5.5], [4.4], [2.2, 3.3, 0.0, -10]], make())
    assert ak.min(a, axis=0).to_list() == [1.1, 3.3, 0, -10]
    assert ak.argmin(a, axis=0).to_list() == [0, 2, 2, 2]


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("case", range(len(AXIS1_CASES)))
def test_jagged_axis1_positions(make, case):
    rows, want = AXIS1_CASES[case]
    a = ak.from_list(rows, make())
    assert ak.min(a, axis=1).to_list() == [[1, 2, 3.3], [1, 2, 3.3]]
    assert ak.argmin(a, axis=1).to_list() == want


@pytest.mark.parametrize("make", BACKENDS)
def test_2d_and_3d(make):
    be = make()
    a2 = ak.from_list([[3.3, 2.2, 5.5, 1.1, 4.4], [4.4, 2.2, 1.1, 3.3, 5.5], [2.2, 1.1, 4.4, 3.3, 5.5]], be)
    assert ak.argmin(a2, axis=0).to_list() == [2, 2, 1, 0, 0]
    assert ak.argmin(a2, axis=1).to_list() == [3, 2, 1]
    a3 = ak.from_list(ARRAY_3D, be)
    by_axis = [[[1, 0, 1, 0, 0], [0, 1, 0, 1, 0], [0, 0, 0, 0, 1]], [[2, 2, 1, 0, 0], [0, 1, 0, 1, 2]],
               [[3, 2, 1], [2, 3, 4]]]
    for axis, want in enumerate(by_axis):
        assert ak.argmin(a3, axis=axis).to_list() == want, axis
        assert ak.argmin(a3, axis=axis - 3).to_list() == want, axis - 3


@pytest.mark.parametrize("make", BACKENDS)
def test_primes_and_missing(make):
    be = make()
    primes = ak.from_list([[[2, 3, 5], [], [7, 11], [13]], [], [[17, 19], [23]]], be, dtype=np.int64)
    assert ak.min(primes, axis=-1).to_list() == [[2, None, 7, 13], [], [17, 23]]
    assert ak.min(primes, axis=-2).to_list() == [[2, 3, 5], [], [17, 19]]
    assert ak.max(primes, axis=-1).to_list() == [[5, None, 11, 13], [], [19, 23]]
    assert ak.max(primes, axis=-2).to_list() == [[13, 11, 5], [], [23, 19]]
    missing = ak.from_list([[0, 1, None], [None, 3], [], None, [4, 5, None, 6]], be, dtype=np.int64)
    assert ak.sum(missing, axis=-2).to_list() == [4, 9, 0, 6]
    assert ak.min(missing, axis=-2).to_list() == [0, 1, None, 6]


def random_3level(seed, n_outer=60):
    rng = np.random.default_rng(seed)
    return [[[int(x) for x in rng.integers(-20, 20, int(rng.poisson(3)))] for _ in range(int(rng.poisson(3)))]
            for _ in range(n_outer)]


def closed_form_axis(rows, axis, op):
    """Reduce a three-level list at axis 0 or 1 (positions aligned from the
    left, missing positions skipped; arg-ops count the rows before)."""
    def combine(groups):  # groups: list of (position, list)
        width = max((len(v) for _, v in groups), default=0)
        out = []
        for j in range(width):
            cands = [(p, v[j]) for p, v in groups if j < len(v)]
            vals = [v for _, v in cands]
            if op == "sum":
                out.append(sum(vals))
            elif op == "max":
                out.append(max(vals))
            elif op == "count":
                out.append(len(vals))
            elif op == "argmax":
                best = cands[0]
                for c in cands[1:]:
                    if c[1] > best[1]:
                        best = c
                out.append(best[0])
        return out

    if axis == 1:
        return [combine(list(enumerate(row))) for row in rows]
    # axis 0: element-wise across the outer rows, per sublist position
    width = max((len(r) for r in rows), default=0)
    res = []
    for i in range(width):
        groups = [(p, r[i]) for p, r in enumerate(rows) if i < len(r)]
        # positions within the sublist-i group, as the regrouping numbers them
        res.append(combine([(k, v) for k, (_, v) in enumerate(groups)]))
    return res


@pytest.mark.parametrize("make", BACKENDS)
@pytest.mark.parametrize("axis", [0, 1])
def test_three_levels_against_closed_form(make, axis):
    be = make()
    rows = random_3level(5)
    a = ak.from_list(rows, be, dtype=np.int64)
    for op in ("sum", "max", "count", "argmax"):
        got = getattr(ak, op)(a, axis=axis).to_list()
        assert got == closed_form_axis(rows, axis, op), (op, axis)


# ---- kernel level, GPU vs oracle ----------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("longest", [60, 600])
def test_nonlocal_kernels_large(longest):
    """maxcount <= 256 takes the sort-free regrouping (ballot counts per
    column), longer rows the stable argsort by column."""
    rng = np.random.default_rng(41)
    length = 60000
    counts = rng.poisson(4, length)
    counts[::501] = rng.integers(20, longest, len(counts[::501]))
    counts[rng.random(length) < 0.1] = 0
    offsets = offsets_from_counts(counts)
    cuts = np.sort(rng.integers(0, length + 1, 999))
    outer = np.concatenate([[0], cuts, [length]]).astype(np.int64)
    outlength = len(outer) - 1
    h, d, he, de = run_both("awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64",
                            [np.zeros(1, np.int64), np.zeros(length + 1, np.int64), offsets, length])
    assert he.str is None and de.str is None and same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    maxcount = int(h[0][0])
    G = outlength * maxcount
    nl = int(offsets[-1])
    h, d, he, de = run_both("awkward_ListOffsetArray_reduce_nonlocal_preparenext_64",
                            [np.zeros(nl, np.int64), np.zeros(G + 1, np.int64), nl, np.zeros(1, np.int64),
                             np.zeros(G, np.int64), G, np.zeros(length + 1, np.int64), offsets, length, outer,
                             outlength, maxcount])
    assert he.str is None and de.str is None
    for k in (0, 1, 3, 4):
        assert same_bits(h[k], d[k]), k
    nextcarry, distincts = h[0], h[4]
    h, d, he, de = run_both("awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64",
                            [np.zeros(maxcount, np.int64), np.zeros(nl, np.int64), np.zeros(nl, np.int64), offsets,
                             length, np.zeros(1, np.int64), outer, outlength, maxcount, nl, nextcarry])
    assert he.str is None and de.str is None
    for k in range(3):
        assert same_bits(h[k], d[k]), k
    h, d, he, de = run_both("awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64",
                            [np.zeros(outlength, np.int64), np.zeros(outlength, np.int64), distincts, G, outlength])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])


@pytest.mark.gpu
def test_nonlocal_reducers_hip_vs_oracle():
    rows = random_3level(6, n_outer=3000)
    out = {}
    for label, make in (("oracle", oracle_backend), ("hip", hip_backend)):
        a = ak.from_list(rows, make(), dtype=np.int64)
        out[label] = {(op, axis): getattr(ak, op)(a, axis=axis).to_list()
                      for op in ("sum", "max", "min", "argmax", "argmin", "count") for axis in (0, 1, 2)}
    assert out["hip"] == out["oracle"]


@pytest.mark.gpu
def test_axis0_long_columns_hip_vs_oracle():
    """axis=0 on 200,000 lists of Poisson(10) lengths: a few columns of up to
    200,000 elements each, which the reducers hand to the long-list queue;
    with 10% of the values missing (IndexedOptionArray) the arg-reductions
    also take the shifts entry points."""
    from awkward_amd import contents as C

    rng = np.random.default_rng(12)
    counts = rng.poisson(10, 200_000)
    off = offsets_from_counts(counts)
    x = rng.integers(-50, 50, int(off[-1])).astype(np.int64)
    index = np.arange(len(x), dtype=np.int64)
    index[rng.random(len(x)) < 0.1] = -1
    out = {}
    for label, make in (("oracle", oracle_backend), ("hip", hip_backend)):
        be = make()
        nl = be.nplike
        plain = ak.from_offsets(off, x, be)
        opt = ak.Array(C.ListOffsetArray(nl.asarray(off), C.IndexedOptionArray(
            nl.asarray(index), C.NumpyArray(nl.asarray(x), be), be), be))
        out[label] = {(op, name): getattr(ak, op)(arr, axis=0).to_list()
                      for op in ("sum", "max", "argmax", "argmin", "count")
                      for name, arr in (("plain", plain), ("option", opt))}
    for key in out["oracle"]:
        assert out["hip"][key] == out["oracle"][key], key

