"""Reductions over the pairs of ak.combinations(ev, 2), recorded lazily and
fused into one pair-kernel launch (awkward_hip_pair_reduce), and with the
plan's n = 2 / n = 3 combinations outputs into one
awkward_hip_combinations_pairs launch -- checked bit for bit against the same
graph evaluated node by node through the host layer, and against the oracle
where the expression's ops are exact IEEE arithmetic."""

import numpy as np
import pytest

import awkward_amd as ak
from awkward_amd import lazy as L
from test_lazy import events, oracle_backend, same


def mass_like(pairs, legs=("0", "1")):
    p0, p1 = pairs[legs[0]], pairs[legs[1]]
    return 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))


def reductions(x):
    return [x.max(), x.argmax(), x.min(), x.argmin(), x.sum()]


def test_ufuncs_record_nodes_and_evaluate_eagerly():
    off, fields = events(2000, seed=7)
    arr = ak.from_offsets(off, fields, oracle_backend())
    lz = L.lazy(arr)
    expr = np.cosh(lz.eta) - np.cos(lz.phi) * np.float32(2) > lz.pt
    assert "cosh" in expr.visualize() and "compare" in expr.visualize()
    got = expr.compute()
    want = np.cosh(arr.eta) - np.cos(arr.phi) * np.float32(2) > arr.pt
    assert got.to_list() == want.to_list()
    assert np.maximum(lz.pt, lz.eta).compute().to_list() == np.maximum(arr.pt, arr.eta).to_list()


def test_pair_graph_evaluates_eagerly_off_the_hip_backend():
    off, fields = events(3000, seed=8)
    arr = ak.from_offsets(off, fields, oracle_backend())
    pairs = L.lazy(arr).combinations(2)
    outs = reductions(mass_like(pairs)) + [pairs]
    plan = L.Plan(outs)
    assert plan.stats()["fused_launches"] == 0  # not a hip array: every output is eager
    got = plan.run()
    p = ak.combinations(arr, 2)
    m = mass_like(p)
    want = [ak.max(m, axis=-1), ak.argmax(m, axis=-1), ak.min(m, axis=-1), ak.argmin(m, axis=-1),
            ak.sum(m, axis=-1)]
    for g, w in zip(got, want):
        assert same(g, w)
    assert got[-1].to_list() == p.to_list()


def heavy_events(n, seed):
    """Events with a few long lists (their workgroups hold more than 2048
    pairs: the per-list path), empty and one-element lists among them."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(4, n)
    counts[rng.choice(n, 12, replace=False)] = rng.integers(65, 300, 12)
    counts[:3] = [0, 1, 2]
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    m = int(off[-1])
    fields = {"pt": rng.exponential(20, m).astype(np.float32), "eta": rng.normal(0, 1.5, m).astype(np.float32),
              "phi": rng.uniform(-np.pi, np.pi, m).astype(np.float32)}
    return off, fields


@pytest.mark.gpu
@pytest.mark.parametrize("make", ["poisson", "heavy"])
def test_fused_pair_reductions_equal_eager(make):
    from awkward_amd.backend import HipBackend

    off, fields = events(40000, seed=9) if make == "poisson" else heavy_events(20000, seed=10)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    pairs = ev.combinations(2)
    outs = reductions(mass_like(pairs)) + [pairs, ev.combinations(3)]
    plan = L.Plan(outs)
    stats = plan.stats()
    assert stats["fused_reductions"] == 5 and stats["fused_launches"] == 1 and stats["eager_outputs"] == 0
    fused = plan.run()
    eager = L.compute(*outs, fuse=False)
    for f, e in zip(fused[:4], eager[:4]):
        assert same(f, e)
    if make == "poisson":
        assert same(fused[4], eager[4])
    else:
        # the pair kernel sums left to right, as the reference's CPU kernel
        # does; the eager reduction of a list of thousands of values adds
        # in a tree, so the two sums agree to float32 rounding only: each is
        # within (n - 1) u sum|x| of the exact sum (recursive summation's
        # bound, u = 2^-24; the tree's log2(n) u is smaller), and the values
        # are >= 0, so sum|x| is the sum itself
        f, e = (x.layout.nplike.to_numpy(x.layout.data).astype(np.float64) for x in (fused[4], eager[4]))
        counts = np.diff(off)
        npairs = counts * (counts - 1) // 2
        assert np.all(np.abs(f - e) <= 2 * np.maximum(npairs - 1, 0) * 2.0**-24 * np.abs(e))
        assert np.array_equal(f[npairs <= 32], e[npairs <= 32])  # short lists: both add left to right
    for f, e in zip(fused[5:], eager[5:]):
        lf, le = f.layout, e.layout
        assert np.array_equal(lf.nplike.to_numpy(lf.offsets), le.nplike.to_numpy(le.offsets))
        for cf, ce in zip(lf.content.contents, le.content.contents):
            assert np.array_equal(cf.nplike.to_numpy(cf.index), ce.nplike.to_numpy(ce.index))
    assert fused[5].to_list()[:200] == eager[5].to_list()[:200]


@pytest.mark.gpu
def test_exact_arithmetic_equals_the_oracle():
    """+, -, * in float32 round the same on every IEEE machine: the fused
    reductions equal the oracle's bit for bit (max, argmax, min, argmin and
    the left-to-right sums)."""
    from awkward_amd.backend import HipBackend

    off, fields = heavy_events(8000, seed=11)
    outs = {}
    for name, be in (("hip", HipBackend()), ("oracle", oracle_backend())):
        pairs = L.lazy(ak.from_offsets(off, fields, be)).combinations(2)
        p0, p1 = pairs["0"], pairs["1"]
        expr = p0.pt * p1.pt - 3 * (p0.eta + p1.eta) * p0.phi
        plan = L.Plan(reductions(expr))
        if name == "hip":
            assert plan.stats()["fused_launches"] == 1
        outs[name] = plan.run()
    for g, w in zip(outs["hip"], outs["oracle"]):
        assert same(g, w)


@pytest.mark.gpu
def test_numbers_float64_and_named_legs():
    """A list of float64 numbers, legs named by `fields`, a constant first
    and a comparison inside the expression."""
    from awkward_amd.backend import HipBackend

    rng = np.random.default_rng(12)
    counts = rng.poisson(6, 30000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    x = rng.standard_normal(int(off[-1]))
    results = {}
    for fuse in (True, False):
        pairs = L.lazy(ak.from_offsets(off, x, HipBackend())).combinations(2, fields=("a", "b"))
        expr = 0.5 * pairs.a - pairs.b * (pairs.a > pairs.b)
        plan = L.Plan(reductions(expr), fuse=fuse)
        if fuse:
            assert plan.stats()["fused_launches"] == 1
        results[fuse] = plan.run()
    for f, e in zip(results[True], results[False]):
        assert same(f, e)


@pytest.mark.gpu
def test_pair_plan_graph_replays():
    import torch

    from awkward_amd.backend import HipBackend

    off, fields = events(30000, seed=13)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    pairs = ev.combinations(2)
    expr = mass_like(pairs)
    triples = ev.combinations(3)
    plan = L.Plan([expr.max(), expr.argmax(), pairs, triples])
    want = L.compute(expr.max(), expr.argmax(), pairs, triples, fuse=False)
    graph = plan.graph()
    first = None
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        unit = plan._unit(next(iter(plan.units)))
        got = dict(unit.out)
        for n in (2, 3):
            for c, carry in enumerate(unit.tuples[n].carry):
                got[f"carry{n}{c}"] = carry[:unit.tuples[n].total]
            got[f"offsets{n}"] = unit.tuples[n].tooffsets
        snap = {k: v.clone() for k, v in got.items()}
        if first is None:
            for n, w in ((2, want[2]), (3, want[3])):
                assert torch.equal(snap[f"offsets{n}"], w.layout.offsets)
                for c, content in enumerate(w.layout.content.contents):
                    assert torch.equal(snap[f"carry{n}{c}"], content.index)
        if first is None:
            first = snap
            mx = want[0].layout
            assert torch.equal(snap["max"], mx.content.data) and torch.equal(snap["argmax"], want[1].layout.content.data)
        else:
            assert all(torch.equal(first[k], snap[k]) for k in first)
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays


@pytest.mark.gpu
def test_combinations_outputs_alone_and_with_replacement():
    """Carries only (no reductions) in one launch for n = 2 and 3; n = 3 with
    replacement and n = 4 keep their own decodes."""
    from awkward_amd.backend import HipBackend

    off, fields = heavy_events(6000, seed=14)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    outs = [ev.combinations(2), ev.combinations(3), ev.combinations(3, replacement=True), ev.combinations(4)]
    plan = L.Plan(outs)
    assert len(plan.units) == 1 and plan.stats()["fused_launches"] == 3
    for f, e in zip(plan.run(), L.compute(*outs, fuse=False)):
        lf, le = f.layout, e.layout
        assert np.array_equal(lf.nplike.to_numpy(lf.offsets), le.nplike.to_numpy(le.offsets))
        for cf, ce in zip(lf.content.contents, le.content.contents):
            assert np.array_equal(cf.nplike.to_numpy(cf.index), ce.nplike.to_numpy(ce.index))


def long_events(n, seed):
    """Poisson(4) events with a few lists longer than a wave's stage (400 to
    900 elements: their pairs are spread over the wave with the leaves read
    from memory) and some of 40-300 elements (one staged list over the wave)."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(4, n)
    counts[rng.choice(n, 4, replace=False)] = rng.integers(400, 900, 4)
    counts[rng.choice(n, 10, replace=False)] = rng.integers(40, 300, 10)
    counts[:4] = [0, 1, 2, 3]
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    m = int(off[-1])
    fields = {"pt": rng.exponential(20, m).astype(np.float32), "eta": rng.normal(0, 1.5, m).astype(np.float32),
              "phi": rng.uniform(-np.pi, np.pi, m).astype(np.float32)}
    return off, fields


def special_events(n, seed):
    """Small integer pt (ties everywhere), signed zeros, NaN and +-inf, lists
    whose first element is NaN (every pair (0, j) is NaN, the first pair
    among them) and lists of NaN only."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(5, n)
    counts[rng.choice(n, 6, replace=False)] = rng.integers(40, 120, 6)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    m = int(off[-1])
    pt = rng.integers(-2, 3, m).astype(np.float32)
    pt[rng.random(m) < 0.05] = -0.0
    pt[rng.random(m) < 0.03] = np.nan
    pt[rng.random(m) < 0.01] = np.inf
    pt[rng.random(m) < 0.01] = -np.inf
    starts = off[:-1][(counts > 1) & (rng.random(n) < 0.05)]
    pt[starts] = np.nan
    for e in np.flatnonzero((counts > 1) & (rng.random(n) < 0.01)):
        pt[off[e]:off[e + 1]] = np.nan
    fields = {"pt": pt, "eta": rng.normal(0, 1.5, m).astype(np.float32),
              "phi": rng.uniform(-np.pi, np.pi, m).astype(np.float32)}
    return off, fields


EXTREMA = [("max", "argmax"), ("min", "argmin"), ("argmax",), ("min",), ("max", "argmax", "min", "argmin")]


def extrema(x, names):
    return [getattr(x, name)() for name in names]


@pytest.mark.gpu
@pytest.mark.parametrize("names", EXTREMA, ids=["-".join(n) for n in EXTREMA])
@pytest.mark.parametrize("make", ["poisson", "heavy", "long", "special"])
def test_pair_extrema_equal_eager(make, names):
    """Extrema without a sum: one wave per 64 lists, each lane reducing a run
    of pairs and the runs' partials merged per list -- bit for bit the
    elementwise form's ak.max / ak.argmax / ak.min / ak.argmin, NaN first
    pairs, all-NaN lists, ties and signed zeros included."""
    from awkward_amd.backend import HipBackend

    off, fields = {"poisson": lambda: events(30000, seed=21), "heavy": lambda: heavy_events(20000, seed=22),
                   "long": lambda: long_events(6000, seed=23), "special": lambda: special_events(20000, seed=24)}[make]()
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    pairs = ev.combinations(2)
    exprs = [mass_like(pairs)]
    if make == "special":
        exprs.append(pairs["0"].pt * pairs["1"].pt)
    for expr in exprs:
        outs = extrema(expr, names)
        plan = L.Plan(outs)
        assert plan.stats()["fused_launches"] == 1
        for f, e in zip(plan.run(), L.compute(*outs, fuse=False)):
            assert same(f, e)


@pytest.mark.gpu
def test_pair_extrema_exact_arithmetic_equal_the_oracle():
    """Products and sums of float32 round alike everywhere: the fused extrema
    equal the oracle's bit for bit, including the special values."""
    from awkward_amd.backend import HipBackend

    off, fields = special_events(6000, seed=25)
    outs = {}
    for name, be in (("hip", HipBackend()), ("oracle", oracle_backend())):
        pairs = L.lazy(ak.from_offsets(off, fields, be)).combinations(2)
        p0, p1 = pairs["0"], pairs["1"]
        expr = p0.pt * p1.pt - 3 * (p0.eta + p1.eta) * p0.phi
        plan = L.Plan(extrema(expr, ("max", "argmax", "min", "argmin")))
        if name == "hip":
            assert plan.stats()["fused_launches"] == 1
        outs[name] = plan.run()
    for g, w in zip(outs["hip"], outs["oracle"]):
        assert same(g, w)


@pytest.mark.gpu
def test_pair_extrema_with_the_combinations():
    """The extrema in the same entry point call as the n = 2 and n = 3
    carries (the carries, then the extrema's launch)."""
    from awkward_amd.backend import HipBackend

    off, fields = long_events(5000, seed=26)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    pairs = ev.combinations(2)
    expr = mass_like(pairs)
    outs = [expr.max(), expr.argmin(), pairs, ev.combinations(3)]
    plan = L.Plan(outs)
    assert plan.stats()["fused_launches"] == 1
    fused, eager = plan.run(), L.compute(*outs, fuse=False)
    for f, e in zip(fused[:2], eager[:2]):
        assert same(f, e)
    for f, e in zip(fused[2:], eager[2:]):
        lf, le = f.layout, e.layout
        assert np.array_equal(lf.nplike.to_numpy(lf.offsets), le.nplike.to_numpy(le.offsets))
        for cf, ce in zip(lf.content.contents, le.content.contents):
            assert np.array_equal(cf.nplike.to_numpy(cf.index), ce.nplike.to_numpy(ce.index))


@pytest.mark.gpu
def test_pair_extrema_compiled_at_run_time_equal_the_interpreter(monkeypatch):
    """Extrema without a sum run in a kernel compiled at run time for the
    program (one more compiled kernel for a new program shape); with
    AWKWARD_HIP_JIT=0 the interpreter runs them: the same bits."""
    import ctypes

    from awkward_amd._lib import lib
    from awkward_amd.backend import HipBackend

    compiled = lib().awkward_hip_jit_kernels
    compiled.restype = ctypes.c_int
    off, fields = special_events(8000, seed=27)
    results = {}
    for jit in ("1", "0"):
        monkeypatch.setenv("AWKWARD_HIP_JIT", jit)
        before = compiled()
        pairs = L.lazy(ak.from_offsets(off, fields, HipBackend())).combinations(2)
        p0, p1 = pairs["0"], pairs["1"]
        expr = np.sqrt(abs(p0.pt - p1.phi)) * np.float32(1.25) + np.exp(p0.eta - p1.eta)  # a shape of its own
        results[jit] = L.Plan(extrema(expr, ("max", "argmax", "min", "argmin"))).run()
        assert compiled() == before + (1 if jit == "1" else 0)
    for a, b in zip(results["1"], results["0"]):
        assert same(a, b)


@pytest.mark.gpu
def test_pair_program_with_integer_legs(monkeypatch):
    """int8 legs multiplied in int8 (wrapping), cast and mixed with float32
    -- the ops the compiled kernel takes through the interpreter's own op
    code next to ones it writes out: compiled, interpreted and the oracle
    agree bit for bit (exact IEEE arithmetic)."""
    from awkward_amd.backend import HipBackend

    rng = np.random.default_rng(28)
    counts = rng.poisson(5, 12000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    m = int(off[-1])
    fields = {"pt": rng.integers(-50, 50, m).astype(np.float32),
              "charge": rng.integers(-128, 128, m).astype(np.int8)}
    results = {}
    for name, be, jit in (("jit", HipBackend(), "1"), ("interpreted", HipBackend(), "0"),
                          ("oracle", oracle_backend(), "1")):
        monkeypatch.setenv("AWKWARD_HIP_JIT", jit)
        pairs = L.lazy(ak.from_offsets(off, fields, be)).combinations(2)
        p0, p1 = pairs["0"], pairs["1"]
        expr = (p0.charge * p1.charge) * p0.pt - p1.pt
        plan = L.Plan(reductions(expr))
        if name != "oracle":
            assert plan.stats()["fused_launches"] == 1
        results[name] = plan.run()
    for a, b, c in zip(results["jit"], results["interpreted"], results["oracle"]):
        assert same(a, b) and same(a, c)
