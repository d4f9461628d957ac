"""Elementwise ufuncs on the hip backend (the expression kernels of
csrc/expr.h) against numpy on the same data, and the fused jagged mask
getitem against the oracle backend's reference sequence.

Bars: every integer, boolean and float arithmetic / comparison result
bit-exact against numpy (the same types, resolved by numpy's own
``resolve_dtypes`` with NEP 50 for Python scalars); cosh / cos / sinh / sin /
exp of the device math library within 4 ulp of numpy's float64 result
rounded to the type (third-party math on both sides: parity unpinned for
exact bits); sqrt correctly rounded, so bit-exact.  Masked getitem outputs
bit-exact and errors with the reference's messages.
"""

import numpy as np
import pytest

import awkward_amd as ak

pytestmark = pytest.mark.gpu

INTS = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64"]
FLOATS = ["float32", "float64"]


def hip(fused=True):
    from awkward_amd.backend import HipBackend

    return HipBackend(fused=fused)


def oracle():
    from oracle.backend import OracleBackend

    return OracleBackend()


def data(rng, dtype, n):
    dtype = np.dtype(dtype)
    if dtype == np.bool_:
        return rng.random(n) < 0.5
    if dtype.kind in "iu":
        info = np.iinfo(dtype)
        v = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        v[rng.random(n) < 0.3] = rng.integers(0, 5, int((rng.random(n) < 0.3).sum()) or 1).astype(dtype)[0]
        return v
    v = (rng.standard_normal(n) * 3).astype(dtype)
    k = max(1, n // 40)
    v[rng.integers(0, n, k)] = rng.choice(np.array([np.nan, np.inf, -np.inf, 0.0, -0.0], dtype=dtype), k)
    return v


def lists(x, be, seed=0):
    rng = np.random.default_rng(seed)
    counts = rng.multinomial(len(x), np.ones(max(len(x) // 3, 1)) / max(len(x) // 3, 1))
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return ak.from_offsets(off, x, be), off


def flat(out):
    layout = out.layout
    inner = layout.content if hasattr(layout, "offsets") else layout
    return np.asarray(inner.nplike.to_numpy(inner.data))


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.dtype == b.dtype, (a.dtype, b.dtype)
    if a.dtype.kind == "f":
        w = np.uint32 if a.dtype.itemsize == 4 else np.uint64
        # NaNs: any NaN payload equals any NaN
        nan = np.isnan(a)
        assert (nan == np.isnan(b)).all()
        assert (a[~nan].view(w) == b[~nan].view(w)).all()
    else:
        assert (a == b).all()


BINARY = [np.add, np.subtract, np.multiply, np.greater, np.greater_equal, np.less, np.less_equal, np.equal,
          np.not_equal, np.maximum, np.minimum]


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("t", INTS + FLOATS)
@pytest.mark.parametrize("ufunc", BINARY, ids=lambda u: u.__name__)
def test_binary_same_type(ufunc, t, fused):
    rng = np.random.default_rng(1)
    n = 5000
    x, y = data(rng, t, n), data(rng, t, n)
    with np.errstate(all="ignore"):
        want = ufunc(x, y)
    a, off = lists(x, hip(fused))
    b = ak.from_offsets(off, y, a.backend)
    same(flat(ufunc(a, b)), want)


@pytest.mark.parametrize("t", INTS + ["bool"])
@pytest.mark.parametrize("ufunc", [np.bitwise_and, np.bitwise_or, np.bitwise_xor], ids=lambda u: u.__name__)
def test_bitwise(ufunc, t):
    rng = np.random.default_rng(2)
    x, y = data(rng, t, 3000), data(rng, t, 3000)
    a, off = lists(x, hip())
    b = ak.from_offsets(off, y, a.backend)
    same(flat(ufunc(a, b)), ufunc(x, y))


@pytest.mark.parametrize("t", INTS + FLOATS)
def test_unary_exact(t):
    rng = np.random.default_rng(3)
    x = data(rng, t, 4000)
    a, _ = lists(x, hip())
    with np.errstate(all="ignore"):
        same(flat(abs(a)), np.absolute(x))
        if np.dtype(t).kind != "u":
            same(flat(-a), np.negative(x))
        if np.dtype(t).kind in "iu":
            same(flat(~a), np.invert(x))
        same(flat(np.true_divide(a, 3)), np.true_divide(x, 3))
    if np.dtype(t).kind == "f":
        with np.errstate(all="ignore"):
            same(flat(np.sqrt(a)), np.sqrt(x))


@pytest.mark.parametrize("t", FLOATS)
@pytest.mark.parametrize("ufunc", [np.cosh, np.cos, np.sinh, np.sin, np.exp], ids=lambda u: u.__name__)
def test_transcendental_within_4_ulp(ufunc, t):
    rng = np.random.default_rng(4)
    x = (rng.standard_normal(6000) * 4).astype(t)
    a, _ = lists(x, hip())
    got = flat(ufunc(a))
    assert got.dtype == np.dtype(t)
    with np.errstate(all="ignore"):
        ref = ufunc(x.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(t)).astype(np.float64)
    finite = np.isfinite(ref) & np.isfinite(got)
    assert (np.isfinite(ref) == np.isfinite(got)).mean() > 0.999
    assert (np.abs(got[finite].astype(np.float64) - ref[finite]) <= 4 * ulp[finite]).all()


def test_numpy_promotion_with_python_scalars():
    """NEP 50 (numpy 2): the types numpy picks, not torch's."""
    be = hip()
    i32 = ak.from_offsets(np.array([0, 2, 3]), np.array([3, -4, 7], np.int32), be)
    same(flat(i32 * 1.5), np.array([3, -4, 7], np.int32) * 1.5)  # float64
    same(flat(i32 / 3), np.array([3, -4, 7], np.int32) / 3)  # float64
    big = ak.from_offsets(np.array([0, 1, 2]), np.array([16777217, 16777216], np.int64), be)
    assert flat(big > 16777216.5).tolist() == [True, False]  # compared in float64
    i8 = ak.from_offsets(np.array([0, 2, 3]), np.array([-128, 5, 127], np.int8), be)
    same(flat(abs(i8)), np.absolute(np.array([-128, 5, 127], np.int8)))  # -128 stays -128
    assert flat(i8 > 1000).tolist() == [False, False, False]  # exact, as numpy 2 compares
    assert flat(i8 < 1000).tolist() == [True, True, True]
    assert flat(i8 != -1000).tolist() == [True, True, True]
    with pytest.raises(OverflowError):
        i8 + 1000
    u8 = ak.from_offsets(np.array([0, 2]), np.array([0, 200], np.uint8), be)
    assert flat(u8 > -1).tolist() == [True, True]
    f32 = ak.from_offsets(np.array([0, 2]), np.array([1.5, 2.5], np.float32), be)
    assert flat(f32 * 2).dtype == np.float32 and flat(f32 + 1.5).dtype == np.float32
    same(flat(2 * f32 * f32), 2 * np.array([1.5, 2.5], np.float32) * np.array([1.5, 2.5], np.float32))


def test_mixed_uint64_int64_comparisons():
    be = hip()
    u = np.array([0, 1, 2**63 + 5, 2**64 - 1, 7], np.uint64)
    s = np.array([-1, 1, 2**63 - 1, -(2**63), 8], np.int64)
    a, off = lists(u, be)
    b = ak.from_offsets(off, s, be)
    for f in (np.greater, np.greater_equal, np.less, np.less_equal, np.equal, np.not_equal):
        same(flat(f(a, b)), f(u, s))
        same(flat(f(b, a)), f(s, u))


def test_chain_is_one_lazy_expression():
    """A chain of ufuncs stays unevaluated until used and then equals
    numpy's chain; common subexpressions and gathered operands included."""
    from awkward_amd.elementwise import DeviceExpr

    rng = np.random.default_rng(5)
    n = 10000
    x = (rng.standard_normal(n) * 10).astype(np.float32)
    y = (rng.standard_normal(n) * 10).astype(np.float32)
    be = hip()
    a, off = lists(x, be)
    b = ak.from_offsets(off, y, be)
    e = (a > 1) & ((abs(b) < 2.4) | (a * b > 5.0))
    assert isinstance(e.layout.content.raw, DeviceExpr)
    want = (x > 1) & ((np.abs(y) < 2.4) | (x * y > 5.0))
    same(flat(e), want)
    s = (a - b) * (a - b) + a
    same(flat(s), (x - y) * (x - y) + x)


def test_gathered_operands_pair_expression():
    """The pair expression over ak.combinations legs: the lazy carries are
    read in place (never gathered into temporaries); the arithmetic matches
    numpy's float32 chain bit for bit, the full expression with cosh / cos
    is within 1e-5 of its scale 2 pt1 pt2 (cosh + 1) of numpy's (the
    difference cosh - cos cancels, so a few-ulp difference of the two math
    libraries is relative to that scale), and ak.max over it equals the
    fused pair kernel's bits (the same operations in the same order)."""
    from awkward_amd.fused import PairMax
    from oracle.backend import OracleBackend

    rng = np.random.default_rng(6)
    nev = 3000
    mult = rng.poisson(4, nev).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(mult)]).astype(np.int64)
    N = int(off[-1])
    fields = {"pt": (5 + rng.exponential(20, N)).astype(np.float32),
              "eta": rng.uniform(-2.5, 2.5, N).astype(np.float32),
              "phi": rng.uniform(-np.pi, np.pi, N).astype(np.float32)}

    def expr(be):
        ev = ak.from_offsets(off, fields, be)
        p = ak.combinations(ev, 2)
        p0, p1 = p["0"], p["1"]
        lin = 2 * p0.pt * p1.pt + (p0.eta - p1.eta)
        full = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))
        scale = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) + 1)
        return ev, flat(lin), full, flat(scale)

    ev, hl, hfull, _ = expr(hip())
    _, cl, cfull, scale = expr(OracleBackend())
    same(hl, cl)
    hf, cf = flat(hfull), flat(cfull)
    assert (np.abs(hf.astype(np.float64) - cf) <= 1e-5 * scale).all()
    pm = PairMax(ev).run()
    m = ak.max(hfull, axis=-1, mask_identity=False)
    same(flat(m), pm["max"].cpu().numpy())


@pytest.mark.parametrize("fused", [True, False])
def test_mask_getitem_matches_oracle(fused):
    rng = np.random.default_rng(7)
    nev = 20000
    mult = np.clip(rng.poisson(3, nev), 0, 20).astype(np.int64)
    mult[::97] = rng.integers(30, 200, len(mult[::97]))  # some lists longer than a tile's stage share
    off = np.concatenate([[0], np.cumsum(mult)]).astype(np.int64)
    N = int(off[-1])
    fields = {"pt": (5 + rng.exponential(20, N)).astype(np.float32),
              "eta": rng.uniform(-2.5, 2.5, N).astype(np.float32)}

    def run(be):
        ev = ak.from_offsets(off, fields, be)
        sel = ev[(ev.pt > 25) & (abs(ev.eta) < 2.4)]
        sel2 = ev[ev.pt > 40][:, 1:]
        return sel.pt.to_list(), ak.num(sel).to_list(), sel2.eta.to_list()

    assert run(hip(fused)) == run(oracle())


def test_mask_getitem_long_lists_unstaged():
    """Tiles whose element span exceeds the stage evaluate the predicate
    from global memory."""
    rng = np.random.default_rng(8)
    counts = np.array([20000, 0, 3, 15000, 1] * 40, dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    x = rng.standard_normal(int(off[-1])).astype(np.float64)

    def run(be):
        a = ak.from_offsets(off, x, be)
        return a[(a > 0.5) | (a < -2)].to_list()

    assert run(hip()) == run(oracle())


def test_mask_getitem_errors_match_reference():
    """A mask list longer than the indexed list with a true value beyond it
    is jagged_apply's "index out of range" (identity = the list, attempt =
    the local index); a mask with another outer length cannot fit."""
    x = np.arange(10, dtype=np.float64)
    off = np.array([0, 3, 5, 10], dtype=np.int64)
    moff = np.array([0, 3, 7, 12], dtype=np.int64)
    m = np.zeros(12, dtype=np.bool_)
    m[[0, 4, 6, 11]] = True  # list 1: local index 3 >= 2
    msgs = []
    for be in (hip(), oracle()):
        a = ak.from_offsets(off, x, be)
        mask = ak.from_offsets(moff, m, be)
        with pytest.raises(IndexError) as info:
            a[mask]
        msgs.append(str(info.value).split(" (in compiled code")[0])
        with pytest.raises(IndexError, match="cannot fit jagged slice"):
            a[ak.from_offsets(np.array([0, 3]), m[:3], be)]
    assert msgs[0] == msgs[1] == "index out of range while attempting to get index 3"


CUTS = [
    lambda e: e.a > 0.5,
    lambda e: (e.a > 0.25) & (abs(e.b) < 1.5),
    lambda e: (e.a >= 0.0) & (e.a <= 2.0) & (e.b != 0.0) & (abs(e.c) > 0.1),
    lambda e: (e.a == 0.0) & (e.b < np.float32(1.25)),
    lambda e: np.logical_and(e.c < 3, abs(e.a) >= 0.5),
    lambda e: (abs(e.a) < 1) & (abs(e.a) > -1) & (e.b > -5) & (e.b < 5) & (e.c > -5) & (e.c < 5) & (e.a != 7)
    & (e.b != 7),
    # not a conjunction of comparisons: the expression interpreter's path
    lambda e: (e.a > 0.5) | (e.b < -1),
    lambda e: ~(e.a > 0.5) & (e.b < 1),
    lambda e: (e.a + e.b) > 0.5,
]


@pytest.mark.parametrize("k", range(len(CUTS)))
def test_mask_getitem_cuts_match_oracle(k):
    """Selection cuts -- conjunctions of float32 comparisons with constants,
    through abs or not, up to eight terms -- take the cut kernel; other
    predicates the interpreter.  Both against the oracle backend's numpy
    evaluation, with NaN, infinities and signed zeros in the columns and
    lengths that end inside a 64-element word and a wave's group."""
    rng = np.random.default_rng(100 + k)
    nev = 7001
    mult = np.clip(rng.poisson(3, nev), 0, 20).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(mult)]).astype(np.int64)
    N = int(off[-1])
    fields = {name: data(rng, "float32", N) for name in "abc"}
    fields["a"][rng.integers(0, N, N // 50)] = 0.0

    def run(be):
        ev = ak.from_offsets(off, fields, be)
        sel = ev[CUTS[k](ev)]
        return sel.a.to_list(), sel.c.to_list(), ak.num(sel).to_list()

    got, want = run(hip()), run(oracle())
    assert got[2] == want[2]
    for g, w in zip(got[:2], want[:2]):
        same(np.array([x for lst in g for x in lst], np.float32), np.array([x for lst in w for x in lst], np.float32))


@pytest.mark.parametrize("slc", [slice(1, None), slice(None, -1), slice(None, None, 2), slice(None, None, -1),
                                 slice(-2, None), slice(1, 3), slice(5, 1, -2), slice(None, 0)])
def test_mask_then_range_matches_oracle(slc):
    """A range slice of a masked getitem's result (offsets made by the
    library's own scan, from 0 to its content's length: the carry is
    allocated to that bound and trimmed, with no carrylength pass), then
    reductions of one field twice (its gather made once), against the
    oracle backend's reference sequence."""
    rng = np.random.default_rng(11)
    nev = 5003
    mult = np.clip(rng.poisson(4, nev), 0, 20).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(mult)]).astype(np.int64)
    N = int(off[-1])
    fields = {"pt": (5 + rng.exponential(20, N)).astype(np.float32), "eta": rng.uniform(-3, 3, N).astype(np.float32)}

    def run(be):
        ev = ak.from_offsets(off, fields, be)
        sel = ev[(ev.pt > 15) & (abs(ev.eta) < 2.4)][:, slc]
        return (ak.num(sel).to_list(), ak.sum(sel.pt, axis=-1).to_list(), ak.argmax(sel.pt, axis=-1).to_list(),
                ak.max(sel.eta, axis=-1).to_list(), sel.eta.to_list())

    assert run(hip()) == run(oracle())
