"""libawkward_hip.so against the oracle on seeded inputs, every family.

Bars: integer, index, mask, offsets, argmin/argmax, max/min and sort results
bit-exact; float sums bit-exact where each list is reduced in the serial order
(lists whose elements fit the workgroup's LDS tile) and otherwise within
|gpu - cpu| <= 2 * n * eps * sum(|x|) per list of n elements (eps of the
output type): the sum of the two orders' worst-case rounding bounds.
"""

import numpy as np
import pytest

from parity import mixed_counts, offsets_from_counts, run_both, same_bits

pytestmark = pytest.mark.gpu

INT_TYPES = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64"]
ALL_TYPES = INT_TYPES + ["float32", "float64"]
NP = {t: np.dtype(t) for t in ALL_TYPES}


def values(rng, dtype, n, special=True):
    dtype = np.dtype(dtype)
    if dtype.kind in "iu":
        info = np.iinfo(dtype)
        v = rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)
        # many ties
        small = rng.random(n) < 0.5
        v[small] = rng.integers(0, 4, size=int(small.sum())).astype(dtype)
        return v
    v = rng.standard_normal(n).astype(dtype)
    if special and n:
        k = max(1, n // 50)
        idx = rng.integers(0, n, size=k)
        v[idx] = rng.choice(np.array([np.nan, 0.0, -0.0, np.inf, -np.inf], dtype=dtype), size=k)
        v[rng.random(n) < 0.2] = rng.choice(np.array([1.0, -1.0, 0.0, -0.0], dtype=dtype), size=1)
    return v


def layouts(rng):
    yield "short", mixed_counts(rng, 3000, lam=4)
    yield "mixed", mixed_counts(rng, 2000, long_every=397, medium_every=13)


@pytest.mark.parametrize("out,inp", [
    ("int64", "int8"), ("uint64", "uint8"), ("int64", "int16"), ("uint64", "uint16"), ("int64", "int32"),
    ("uint64", "uint32"), ("int64", "int64"), ("uint64", "uint64"), ("float32", "float32"), ("float64", "float64"),
    ("int32", "int8"), ("uint32", "uint8"), ("int32", "int16"), ("uint32", "uint16"), ("int32", "int32"),
    ("uint32", "uint32"), ("float64", "int8"), ("float64", "uint8"), ("float64", "int16"), ("float64", "uint16"),
    ("float64", "int32"), ("float64", "uint32"), ("float64", "int64"), ("float64", "uint64"),
    ("float64", "float32"),
])
def test_reduce_sum(out, inp):
    rng = np.random.default_rng(11)
    for label, counts in layouts(rng):
        off = offsets_from_counts(counts)
        x = values(rng, inp, int(off[-1]), special=False)
        name = f"awkward_reduce_sum_{out}_{inp}_64"
        h, d, he, de = run_both(name, [np.zeros(len(counts), out), x, off, len(counts)])
        assert he.str is None and de.str is None
        if np.dtype(out).kind != "f":
            assert same_bits(h[0], d[0]), (name, label)
            continue
        if label == "short":
            assert same_bits(h[0], d[0]), (name, label)  # serial order reproduced
        eps = np.finfo(np.dtype(out)).eps
        absum = np.array([np.abs(x[a:b].astype(np.float64)).sum() for a, b in zip(off[:-1], off[1:])])
        bound = 2 * counts * eps * absum
        assert (np.abs(h[0].astype(np.float64) - d[0].astype(np.float64)) <= bound).all()


@pytest.mark.parametrize("out,inp", [
    ("int64", "int8"), ("uint64", "uint8"), ("int64", "int16"), ("uint64", "uint16"), ("int64", "int32"),
    ("uint64", "uint32"), ("int64", "int64"), ("uint64", "uint64"), ("float32", "float32"), ("float64", "float64"),
    ("int32", "int8"), ("uint32", "uint8"), ("int32", "int16"), ("uint32", "uint16"), ("int32", "int32"),
    ("uint32", "uint32"),
])
def test_reduce_prod(out, inp):
    """Integer products wrap bit-exactly; float products (values of magnitude
    0.9-1.1, so long lists neither overflow nor underflow) are bit-exact on
    serial tiles and otherwise within 2 * n * eps * |product| (each order's
    relative rounding bound is n * eps)."""
    rng = np.random.default_rng(13)
    for label, counts in layouts(rng):
        off = offsets_from_counts(counts)
        N = int(off[-1])
        if np.dtype(inp).kind == "f":
            x = (rng.uniform(0.9, 1.1, N) * rng.choice([-1.0, 1.0], N)).astype(inp)
        else:
            x = values(rng, inp, N)
        name = f"awkward_reduce_prod_{out}_{inp}_64"
        h, d, he, de = run_both(name, [np.zeros(len(counts), out), x, off, len(counts)])
        assert he.str is None and de.str is None
        if np.dtype(out).kind != "f" or label == "short":
            assert same_bits(h[0], d[0]), (name, label)
            continue
        eps = np.finfo(np.dtype(out)).eps
        bound = 2 * counts * eps * np.abs(h[0].astype(np.float64))
        assert (np.abs(h[0].astype(np.float64) - d[0].astype(np.float64)) <= bound).all()


@pytest.mark.parametrize("family", ["sum_bool", "prod_bool"])
@pytest.mark.parametrize("t", ["bool"] + ALL_TYPES)
def test_reduce_any_all(family, t):
    """awkward_reduce_sum_bool (any nonzero) / prod_bool (all nonzero):
    bit-exact, with zeros, -0.0 and NaN in the data and all-zero /
    all-nonzero lists of every length class."""
    rng = np.random.default_rng(14)
    for _, counts in layouts(rng):
        off = offsets_from_counts(counts)
        N = int(off[-1])
        if t == "bool":
            x = rng.random(N) < 0.5
        else:
            x = values(rng, t, N)
            x[rng.random(N) < 0.5] = 0
        # whole lists of zeros or of nonzeros
        for i in range(0, len(counts), 7):
            x[off[i]:off[i + 1]] = 0
        for i in range(3, len(counts), 11):
            x[off[i]:off[i + 1]] = 1
        name = f"awkward_reduce_{family}_{t}_64"
        h, d, he, de = run_both(name, [np.zeros(len(counts), np.bool_), x, off, len(counts)])
        assert he.str is None and de.str is None
        assert same_bits(h[0], d[0]), name


@pytest.mark.parametrize("out", ["int64", "int32"])
def test_reduce_sum_bool(out):
    rng = np.random.default_rng(12)
    for _, counts in layouts(rng):
        off = offsets_from_counts(counts)
        x = rng.random(int(off[-1])) < 0.3
        h, d, he, de = run_both(f"awkward_reduce_sum_{out}_bool_64", [np.zeros(len(counts), out), x, off, len(counts)])
        assert same_bits(h[0], d[0])


@pytest.mark.parametrize("t", ALL_TYPES)
@pytest.mark.parametrize("op", ["max", "min"])
def test_reduce_extremum(op, t):
    rng = np.random.default_rng(13)
    dtype = np.dtype(t)
    identities = [None]
    if dtype.kind == "f":
        identities = [-np.inf if op == "max" else np.inf, 0.0]
    for label, counts in layouts(rng):
        off = offsets_from_counts(counts)
        x = values(rng, t, int(off[-1]))
        for ident in identities:
            if ident is None:
                info = np.iinfo(dtype)
                ident = int(info.min if op == "max" else info.max)
            name = f"awkward_reduce_{op}_{t}_{t}_64"
            h, d, _, _ = run_both(name, [np.zeros(len(counts), t), x, off, len(counts), ident])
            assert same_bits(h[0], d[0]), (name, label, ident)


@pytest.mark.parametrize("t", ALL_TYPES)
@pytest.mark.parametrize("op", ["argmax", "argmin"])
def test_reduce_arg(op, t):
    rng = np.random.default_rng(14)
    for label, counts in layouts(rng):
        off = offsets_from_counts(counts)
        x = values(rng, t, int(off[-1]))
        if np.dtype(t).kind == "f":
            x[off[:-1][counts > 0][::7]] = np.nan  # NaN first elements
        name = f"awkward_reduce_{op}_{t}_64"
        h, d, _, _ = run_both(name, [np.zeros(len(counts), np.int64), x, off, off[:-1].copy(), len(counts)])
        assert same_bits(h[0], d[0]), (name, label)


def test_count_adjust_mask():
    rng = np.random.default_rng(15)
    counts = mixed_counts(rng, 5000)
    off = offsets_from_counts(counts)
    n = len(counts)
    h, d, _, _ = run_both("awkward_reduce_count_64", [np.zeros(n, np.int64), off, n])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both("awkward_NumpyArray_reduce_mask_ByteMaskedArray_64", [np.zeros(n, np.int8), off, n])
    assert same_bits(h[0], d[0])
    toptr = np.where(counts > 0, off[:-1] + rng.integers(0, 3, n), -1).astype(np.int64)
    h, d, _, _ = run_both("awkward_NumpyArray_reduce_adjust_starts_64", [toptr, n, off, off[:-1].copy()])
    assert same_bits(h[0], d[0])


@pytest.mark.parametrize("c", ["int32", "uint32", "int64"])
def test_localindex(c):
    rng = np.random.default_rng(16)
    counts = mixed_counts(rng, 4000, long_every=1000)
    off = offsets_from_counts(counts).astype(c)
    p = {"int32": "32", "uint32": "U32", "int64": "64"}[c]
    h, d, _, _ = run_both(f"awkward_ListArray{p}_localindex_64", [np.zeros(int(off[-1]), np.int64), off, len(counts)])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both("awkward_localindex_64", [np.zeros(1000, np.int64), 1000])
    assert same_bits(h[0], d[0])


P = {"int32": "32", "uint32": "U32", "int64": "64"}


@pytest.mark.parametrize("c", ["int32", "uint32", "int64"])
def test_jagged_getitem(c):
    rng = np.random.default_rng(17)
    n = 6000
    counts = mixed_counts(rng, n, lam=3)
    off = offsets_from_counts(counts)
    # a jagged index: per list some positions, negative ones included
    sc = rng.integers(0, 4, n) * (counts > 0)
    soff = offsets_from_counts(sc)
    idx = np.concatenate([rng.integers(-c_, c_, k) if k else np.zeros(0, np.int64)
                          for c_, k in zip(counts, sc)]).astype(np.int64)
    h, d, _, _ = run_both("awkward_RegularArray_getitem_jagged_expand_64",
                          [np.zeros(n, np.int64), np.zeros(n, np.int64), soff, n, 1])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    h, d, _, _ = run_both("awkward_ListArray_getitem_jagged_carrylen_64", [np.zeros(1, np.int64), soff[:-1].copy(),
                                                                           soff[1:].copy(), n])
    assert same_bits(h[0], d[0])
    fs, fe = off[:-1].astype(c), off[1:].astype(c)
    args = [np.zeros(n + 1, np.int64), np.zeros(len(idx), np.int64), soff[:-1].copy(), soff[1:].copy(), n, idx,
            len(idx), fs, fe, int(off[-1])]
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_jagged_apply_64", args)
    assert he.str is None and de.str is None
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    # errors: the first failing list in loop order, same message/identity/attempt
    bad = idx.copy()
    first_nonempty = [i for i in range(n) if sc[i] > 0]
    i1, i2 = first_nonempty[50], first_nonempty[20]
    bad[soff[i1]] = 10**6
    bad[soff[i2] + sc[i2] - 1] = -(10**6)
    args[5] = bad
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_jagged_apply_64", args)
    assert he.str is not None and de.str is not None
    assert (he.str, he.id, he.attempt) == (de.str, de.id, de.attempt)
    # descend
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_jagged_descend_64",
                            [np.zeros(n + 1, np.int64), off[:-1].copy(), off[1:].copy(), n, fs, fe])
    assert he.str is None and de.str is None and same_bits(h[0], d[0])
    fe2 = fe.copy()
    fe2[77] += 1
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_jagged_descend_64",
                            [np.zeros(n + 1, np.int64), off[:-1].copy(), off[1:].copy(), n, fs, fe2])
    assert (he.str, he.id) == (de.str, de.id) and he.str is not None


NONE = 2**63 - 1


@pytest.mark.parametrize("c", ["int32", "uint32", "int64"])
@pytest.mark.parametrize("sl", [(1, NONE, 1), (NONE, NONE, 1), (0, 3, 1), (-2, NONE, 1), (1, -1, 2), (NONE, NONE, -1),
                                (5, 0, -2), (-1, -10, -3), (NONE, 2, -1), (3, 100, 7)])
def test_next_range(c, sl):
    rng = np.random.default_rng(18)
    n = 5000
    counts = mixed_counts(rng, n, lam=5)
    off = offsets_from_counts(counts)
    fs, fe = off[:-1].astype(c), off[1:].astype(c)
    start, stop, step = sl
    h, d, _, _ = run_both(f"awkward_ListArray{P[c]}_getitem_next_range_carrylength",
                          [np.zeros(1, np.int64), fs, fe, n, start, stop, step])
    assert same_bits(h[0], d[0])
    k = int(h[0][0])
    h, d, _, _ = run_both(f"awkward_ListArray{P[c]}_getitem_next_range_64",
                          [np.zeros(n + 1, c), np.zeros(max(k, 1), np.int64), fs, fe, n, start, stop, step])
    assert same_bits(h[0], d[0]) and same_bits(h[1][:k], d[1][:k])
    h, d, _, _ = run_both(f"awkward_ListArray{P[c]}_getitem_next_range_counts_64",
                          [np.zeros(1, np.int64), off.astype(c), n])
    assert same_bits(h[0], d[0])


def test_regular_next_range():
    h, d, _, _ = run_both("awkward_RegularArray_getitem_next_range_64", [np.zeros(300, np.int64), 2, 3, 30, 40, 10])
    assert same_bits(h[0], d[0])


@pytest.mark.parametrize("c", ["int32", "uint32", "int64"])
def test_next_at_compact_nextcarry(c):
    rng = np.random.default_rng(19)
    n = 4000
    counts = rng.integers(2, 9, n)
    off = offsets_from_counts(counts)
    fs, fe = off[:-1].astype(c), off[1:].astype(c)
    for at in (0, -1, 1, -2):
        h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_next_at_64", [np.zeros(n, np.int64), fs, fe, n, at])
        assert he.str is None and de.str is None and same_bits(h[0], d[0])
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_getitem_next_at_64", [np.zeros(n, np.int64), fs, fe, n, 3])
    assert (he.str, he.id, he.attempt) == (de.str, de.id, de.attempt) and he.str is not None
    h, d, he, de = run_both(f"awkward_ListArray{P[c]}_compact_offsets_64", [np.zeros(n + 1, np.int64), fs, fe, n])
    assert same_bits(h[0], d[0])
    fe2 = fe.copy()
    fe2[1234] = fs[1234] - 1 if c != "uint32" or fs[1234] > 0 else 0
    if fe2[1234] < fs[1234]:
        h, d, he, de = run_both(f"awkward_ListArray{P[c]}_compact_offsets_64", [np.zeros(n + 1, np.int64), fs, fe2, n])
        assert (he.str, he.id) == (de.str, de.id) and he.str is not None
    lencontent = 1000
    index = rng.integers(0, lencontent, 5000).astype(c)
    h, d, he, de = run_both(f"awkward_IndexedArray{P[c]}_getitem_nextcarry_64",
                            [np.zeros(5000, np.int64), index, 5000, lencontent])
    assert he.str is None and same_bits(h[0], d[0])
    index[3000] = lencontent + 5
    index[4000] = lencontent + 7
    h, d, he, de = run_both(f"awkward_IndexedArray{P[c]}_getitem_nextcarry_64",
                            [np.zeros(5000, np.int64), index, 5000, lencontent])
    assert (he.str, he.id, he.attempt) == (de.str, de.id, de.attempt) and he.str is not None


@pytest.mark.parametrize("c", ["int32", "uint32", "int64"])
@pytest.mark.parametrize("n,replacement", [(1, False), (2, False), (3, False), (4, False), (2, True), (3, True)])
def test_combinations(c, n, replacement):
    rng = np.random.default_rng(20)
    nl = 3000
    counts = np.clip(rng.poisson(4, nl), 0, 30)
    counts[::101] = rng.integers(20, 40, len(counts[::101]))
    off = offsets_from_counts(counts)
    starts, stops = off[:-1].astype(c), off[1:].astype(c)
    h, d, _, _ = run_both(f"awkward_ListArray{P[c]}_combinations_length_64",
                          [np.zeros(1, np.int64), np.zeros(nl + 1, np.int64), n, replacement, starts, stops, nl])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    total = int(h[0][0])
    carry = [np.zeros(total, np.int64) for _ in range(n)]
    h, d, _, _ = run_both(f"awkward_ListArray{P[c]}_combinations_64",
                          [carry, np.zeros(n, np.int64), np.zeros(n, np.int64), n, replacement, starts, stops, nl])
    for k in range(n):
        assert same_bits(h[0][k], d[0][k]), k
    assert same_bits(h[1], d[1])


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("replacement", [False, True])
def test_combinations_skewed(n, replacement):
    """Runs of thousands of lists without tuples between a few very long
    lists: tiles whose lists overflow the decode's offsets window."""
    rng = np.random.default_rng(25)
    counts = np.zeros(30000, dtype=np.int64)
    counts[rng.integers(0, 30000, 20)] = rng.integers(40, 120, 20)
    counts[:5000] = rng.integers(0, n, 5000)  # no tuples without replacement
    off = offsets_from_counts(counts)
    nl = len(counts)
    starts, stops = off[:-1], off[1:]
    h, d, _, _ = run_both("awkward_ListArray64_combinations_length_64",
                          [np.zeros(1, np.int64), np.zeros(nl + 1, np.int64), n, replacement, starts, stops, nl])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    total = int(h[0][0])
    carry = [np.zeros(total, np.int64) for _ in range(n)]
    h, d, _, _ = run_both("awkward_ListArray64_combinations_64",
                          [carry, np.zeros(n, np.int64), np.zeros(n, np.int64), n, replacement, starts, stops, nl])
    for k in range(n):
        assert same_bits(h[0][k], d[0][k]), k
    assert same_bits(h[1], d[1])


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("replacement", [False, True])
def test_combinations_table_edges(n, replacement):
    """List sizes around the decode's colex-table limits (pairs of up to 64
    items, triples of up to 32, counting the n - 1 extra items of
    multisets), on a ListArray whose lists are scattered over the content
    (starts not equal to the previous stops)."""
    rng = np.random.default_rng(26)
    edges = [0, 1, 2, 3, 29, 30, 31, 32, 33, 34, 61, 62, 63, 64, 65, 66]
    sizes = np.array(edges * 40 + list(rng.poisson(5, 3000)), dtype=np.int64)
    rng.shuffle(sizes)
    gaps = rng.integers(0, 7, len(sizes))
    starts = np.cumsum(np.concatenate([[0], (sizes + gaps)[:-1]])).astype(np.int64)
    stops = starts + sizes
    nl = len(sizes)
    h, d, _, _ = run_both("awkward_ListArray64_combinations_length_64",
                          [np.zeros(1, np.int64), np.zeros(nl + 1, np.int64), n, replacement, starts, stops, nl])
    assert same_bits(h[0], d[0]) and same_bits(h[1], d[1])
    total = int(h[0][0])
    carry = [np.zeros(total, np.int64) for _ in range(n)]
    h, d, _, _ = run_both("awkward_ListArray64_combinations_64",
                          [carry, np.zeros(n, np.int64), np.zeros(n, np.int64), n, replacement, starts, stops, nl])
    for k in range(n):
        assert same_bits(h[0][k], d[0][k]), k
    assert same_bits(h[1], d[1])


SORT_TYPES = ["bool"] + ALL_TYPES


def sort_values(rng, t, n):
    if t == "bool":
        return rng.random(n) < 0.5
    return values(rng, t, n)


@pytest.mark.parametrize("t", SORT_TYPES)
@pytest.mark.parametrize("ascending", [True, False])
def test_argsort_sort(t, ascending):
    rng = np.random.default_rng(21)
    counts = np.concatenate([rng.integers(0, 17, 300), rng.integers(17, 65, 100), rng.integers(65, 4097, 10),
                             np.array([4096, 4097, 9000, 20000]), np.array([0, 1, 0])])
    rng.shuffle(counts)
    off = offsets_from_counts(counts)
    x = sort_values(rng, t, int(off[-1]))
    for u in ["int32", "uint32", "uint64", "int64"]:
        name = f"awkward_argsort_{t}_{u}"
        h, d, _, _ = run_both(name, [np.zeros(len(x), np.int64), x, len(x), off.astype(u), len(off), ascending, True])
        assert same_bits(h[0], d[0]), (name, ascending)
    # stable=False: the device always returns the stable order, one valid answer
    h, d, _, _ = run_both(f"awkward_argsort_{t}_int64", [np.zeros(len(x), np.int64), x, len(x), off, len(off),
                                                         ascending, False])
    hs, _, _, _ = run_both(f"awkward_argsort_{t}_int64", [np.zeros(len(x), np.int64), x, len(x), off, len(off),
                                                          ascending, True])
    assert same_bits(hs[0], d[0])
    h, d, _, _ = run_both(f"awkward_sort_{t}", [np.zeros(len(x), x.dtype), x, len(x), off, len(off), ascending, True])
    assert same_bits(h[0], d[0])


def test_argsort_uncovered_and_clamped():
    rng = np.random.default_rng(22)
    x = rng.standard_normal(100)
    off = np.array([10, 20, 20, 55, 150], dtype=np.int64)  # starts past 0, ends past len
    h, d, _, _ = run_both("awkward_argsort_float64_int64", [np.zeros(100, np.int64), x, 100, off, len(off), True, True])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both("awkward_sort_float64", [np.zeros(100, np.float64), x, 100, off, len(off), True, True])
    assert same_bits(h[0][:100], d[0][:100])


def test_local_preparenext():
    rng = np.random.default_rng(23)
    for n in (10000, 300000):  # the second is one queued long segment
        x = rng.integers(0, 50, n).astype(np.int64)
        h, d, _, _ = run_both("awkward_ListOffsetArray_local_preparenext_64", [np.zeros(n, np.int64), x, n])
        assert same_bits(h[0], d[0]), n


@pytest.mark.parametrize("t", ["int32", "int64", "float32", "float64"])
@pytest.mark.parametrize("ascending", [True, False])
def test_long_segments(t, ascending):
    """Segments past 4096 elements go through the queued multi-workgroup
    merge sort: exact multiples of the chunk, one past, and up to 1e5."""
    rng = np.random.default_rng(24)
    counts = np.array([8192, 12288, 4097, 100000, 3, 0, 65537, 5000, 33000, 16384, 8193])
    rng.shuffle(counts)
    off = offsets_from_counts(counts)
    x = values(rng, t, int(off[-1]))
    h, d, _, _ = run_both(f"awkward_argsort_{t}_int64", [np.zeros(len(x), np.int64), x, len(x), off, len(off),
                                                         ascending, True])
    assert same_bits(h[0], d[0])
    h, d, _, _ = run_both(f"awkward_sort_{t}", [np.zeros(len(x), x.dtype), x, len(x), off, len(off), ascending, True])
    assert same_bits(h[0], d[0])


@pytest.mark.parametrize("t", ["bool", "int8", "uint16", "uint32"])
def test_long_segments_narrow_keys(t):
    """Long segments of the types whose order keys travel as 32 bits (every
    type of at most 4 bytes), both directions, with many ties."""
    rng = np.random.default_rng(27)
    counts = np.array([9000, 4097, 70000, 12288, 5])
    off = offsets_from_counts(counts)
    x = sort_values(rng, t, int(off[-1]))
    for ascending in (True, False):
        h, d, _, _ = run_both(f"awkward_argsort_{t}_int64", [np.zeros(len(x), np.int64), x, len(x), off, len(off),
                                                             ascending, True])
        assert same_bits(h[0], d[0]), ascending
        h, d, _, _ = run_both(f"awkward_sort_{t}", [np.zeros(len(x), x.dtype), x, len(x), off, len(off), ascending,
                                                    True])
        assert same_bits(h[0], d[0]), ascending


@pytest.mark.parametrize("variant", range(10))
@pytest.mark.parametrize("t", ["int32", "int64", "float32", "float64"])
def test_tuned_variants(variant, t):
    """The tuned/fused entry points, every tile variant, against the oracle's
    plain kernels followed by adjust_starts / reduce_mask."""
    import torch

    from awkward_amd import dispatch

    rng = np.random.default_rng(30 + variant)
    for label, counts in [("sparse", np.clip(rng.poisson(0.4, 20000), 0, 20)),
                          ("short", mixed_counts(rng, 6000, lam=10)),
                          ("mixed", mixed_counts(rng, 3000, long_every=499, medium_every=7))]:
        off = offsets_from_counts(counts)
        n = len(counts)
        x = values(rng, t, int(off[-1]), special=False)
        d_off = torch.from_numpy(off).cuda()
        d_x = torch.from_numpy(x).cuda()
        # sum
        out_t = {"int32": "int64", "int64": "int64", "float32": "float32", "float64": "float64"}[t]
        h, _, _, _ = run_both(f"awkward_reduce_sum_{out_t}_{t}_64", [np.zeros(n, out_t), x, off, n])
        d_out = torch.empty(n, dtype=getattr(torch, out_t), device="cuda")
        assert dispatch.reduce_sum(d_out, d_x, d_off, n, variant).str is None
        got = d_out.cpu().numpy()
        if out_t.startswith("int") or label != "mixed":
            assert same_bits(h[0], got), ("sum", label, variant)
        # max with mask
        ident = -np.inf if t.startswith("float") else int(np.iinfo(t).min)
        h, _, _, _ = run_both(f"awkward_reduce_max_{t}_{t}_64", [np.zeros(n, t), x, off, n, ident])
        hm, _, _, _ = run_both("awkward_NumpyArray_reduce_mask_ByteMaskedArray_64", [np.zeros(n, np.int8), off, n])
        d_out = torch.empty(n, dtype=getattr(torch, t), device="cuda")
        d_mask = torch.empty(n, dtype=torch.int8, device="cuda")
        assert dispatch.reduce_extremum("max", d_out, d_mask, d_x, d_off, n, ident, variant).str is None
        assert same_bits(h[0], d_out.cpu().numpy()) and same_bits(hm[0], d_mask.cpu().numpy()), ("max", label)
        # argmin, local
        h, _, _, _ = run_both(f"awkward_reduce_argmin_{t}_64", [np.zeros(n, np.int64), x, off, off[:-1].copy(), n])
        want = np.where(h[0] >= 0, h[0] - off[:-1], h[0])
        d_out = torch.empty(n, dtype=torch.int64, device="cuda")
        assert dispatch.reduce_arg("argmin", d_out, d_mask, d_x, d_off, n, variant).str is None
        assert same_bits(want, d_out.cpu().numpy()) and same_bits(hm[0], d_mask.cpu().numpy()), ("argmin", label)


@pytest.mark.parametrize("t", ["int32", "int64", "float32", "float64"])
def test_long_list_queue_and_shifts(t):
    """Variant 10 (lists longer than 4096 elements split into 8192-element
    chunks across the chip, partials combined in order) on a few very long
    lists among short and empty ones, and the arg-reductions' shifts entry
    points (adjust_starts_shifts in the epilogue), against the oracle."""
    import torch

    from awkward_amd import dispatch

    rng = np.random.default_rng(77)
    counts = np.array([300000, 0, 5, 4097, 200000, 1, 0, 8192, 8193, 123457, 4096, 2, 40000], dtype=np.int64)
    off = offsets_from_counts(counts)
    n = len(counts)
    x = values(rng, t, int(off[-1]), special=False)
    if t.startswith("int"):
        x[rng.integers(0, len(x), 64)] = np.iinfo(t).max  # ties of the maximum across chunks
    d_off = torch.from_numpy(off).cuda()
    d_x = torch.from_numpy(x).cuda()
    out_t = {"int32": "int64", "int64": "int64", "float32": "float32", "float64": "float64"}[t]
    h, _, _, _ = run_both(f"awkward_reduce_sum_{out_t}_{t}_64", [np.zeros(n, out_t), x, off, n])
    d_out = torch.empty(n, dtype=getattr(torch, out_t), device="cuda")
    assert dispatch.reduce_sum(d_out, d_x, d_off, n, dispatch.LONG_LISTS).str is None
    got = d_out.cpu().numpy()
    if out_t.startswith("int"):
        assert same_bits(h[0], got)
    else:
        eps = np.finfo(out_t).eps
        absum = np.add.reduceat(np.abs(x.astype(np.float64)), off[:-1]) if len(x) else np.zeros(n)
        absum[counts == 0] = 0
        bound = 2 * counts * eps * absum
        assert np.all(np.abs(got.astype(np.float64) - h[0].astype(np.float64)) <= bound)
    hm, _, _, _ = run_both("awkward_NumpyArray_reduce_mask_ByteMaskedArray_64", [np.zeros(n, np.int8), off, n])
    d_mask = torch.empty(n, dtype=torch.int8, device="cuda")
    for kind in ("max", "min"):
        ident = (-np.inf if kind == "max" else np.inf) if t.startswith("float") else \
            int(np.iinfo(t).min if kind == "max" else np.iinfo(t).max)
        h, _, _, _ = run_both(f"awkward_reduce_{kind}_{t}_{t}_64", [np.zeros(n, t), x, off, n, ident])
        d_val = torch.empty(n, dtype=getattr(torch, t), device="cuda")
        assert dispatch.reduce_extremum(kind, d_val, d_mask, d_x, d_off, n, ident, dispatch.LONG_LISTS).str is None
        assert same_bits(h[0], d_val.cpu().numpy()) and same_bits(hm[0], d_mask.cpu().numpy()), kind
    shifts = np.cumsum(rng.integers(0, 3, len(x))).astype(np.int64)
    starts = off[:-1] + np.cumsum(rng.integers(0, 5, n))  # the lists' starts before compaction
    d_shifts = torch.from_numpy(shifts).cuda()
    d_starts = torch.from_numpy(starts).cuda()
    for kind in ("argmax", "argmin"):
        h, _, _, _ = run_both(f"awkward_reduce_{kind}_{t}_64", [np.zeros(n, np.int64), x, off, off[:-1].copy(), n])
        local = np.where(h[0] >= 0, h[0] - off[:-1], h[0])
        hs, _, _, _ = run_both("awkward_NumpyArray_reduce_adjust_starts_shifts_64",
                               [h[0].copy(), n, np.zeros(len(x), np.int64), starts, shifts])
        for variant in (dispatch.LONG_LISTS, 3):
            d_out = torch.empty(n, dtype=torch.int64, device="cuda")
            assert dispatch.reduce_arg(kind, d_out, d_mask, d_x, d_off, n, variant).str is None
            assert same_bits(local, d_out.cpu().numpy()), (kind, variant)
            assert same_bits(hm[0], d_mask.cpu().numpy()), (kind, variant)
            assert dispatch.reduce_arg(kind, d_out, None, d_x, d_off, n, variant, shifts=d_shifts,
                                       starts=d_starts).str is None
            assert same_bits(hs[0], d_out.cpu().numpy()), (kind, variant, "shifts")



def test_long_list_queue_in_a_graph():
    """Variant 10 reads nothing back to size its queue (a fixed capacity;
    lists that would overflow it stay with their workgroup), so a sum,
    masked max and local argmax over very long lists can be captured:
    every replay -- with eager library work and host synchronisations in
    between -- equals the eager results and the oracle."""
    import torch

    from awkward_amd import dispatch
    from awkward_amd.fused import Graph

    rng = np.random.default_rng(78)
    counts = np.array([250000, 0, 3, 9000, 100001, 1, 8192, 70000], dtype=np.int64)
    off = offsets_from_counts(counts)
    n = len(counts)
    x = rng.standard_normal(int(off[-1]))
    d_off, d_x = torch.from_numpy(off).cuda(), torch.from_numpy(x).cuda()
    outs = {"sum": torch.empty(n, dtype=torch.float64, device="cuda"),
            "max": torch.empty(n, dtype=torch.float64, device="cuda"),
            "mask": torch.empty(n, dtype=torch.int8, device="cuda"),
            "argmax": torch.empty(n, dtype=torch.int64, device="cuda")}

    def step():
        assert dispatch.reduce_sum(outs["sum"], d_x, d_off, n, dispatch.LONG_LISTS).str is None
        assert dispatch.reduce_extremum("max", outs["max"], outs["mask"], d_x, d_off, n, -np.inf,
                                        dispatch.LONG_LISTS).str is None
        assert dispatch.reduce_arg("argmax", outs["argmax"], None, d_x, d_off, n, dispatch.LONG_LISTS).str is None
        return outs

    step()
    torch.cuda.synchronize()
    eager = {k: v.clone() for k, v in outs.items()}
    hmax, _, _, _ = run_both("awkward_reduce_max_float64_float64_64", [np.zeros(n), x, off, n, -np.inf])
    harg, _, _, _ = run_both("awkward_reduce_argmax_float64_64", [np.zeros(n, np.int64), x, off, off[:-1].copy(), n])
    assert same_bits(hmax[0], eager["max"].cpu().numpy())
    assert same_bits(np.where(harg[0] >= 0, harg[0] - off[:-1], harg[0]), eager["argmax"].cpu().numpy())
    g = Graph(step)
    for _ in range(3):
        for v in outs.values():
            v.zero_()
        dispatch.reduce_sum(torch.empty(3, dtype=torch.float64, device="cuda"), d_x, d_off[:4], 3, 3)  # eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in outs:
            assert torch.equal(outs[k], eager[k]), k
