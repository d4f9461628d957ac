"""Fused, graph-capturable forms against the per-kernel path and the oracle."""

import numpy as np
import pytest

import awkward_amd as ak
from parity import same_bits
from test_pipeline import events, flat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cpu():
    from oracle.backend import OracleBackend

    return OracleBackend()


@pytest.fixture(scope="module")
def hip():
    from awkward_amd.backend import hip_backend

    return hip_backend()


TERMS = [("pt", ">", 25.0), ("eta", "abs<", 2.4)]


@pytest.mark.parametrize("sl", [slice(1, None), slice(None), slice(3, None), slice(0, 2), slice(-2, None), slice(1, None, 2)])
def test_event_pipeline_matches_oracle(cpu, hip, sl):
    from awkward_amd.fused import EventPipeline

    off, fields = events(5, 30000, 3, 20)
    ev_c = ak.from_offsets(off, fields, cpu)
    sel = ev_c[(ev_c.pt > 25) & (abs(ev_c.eta) < 2.4)][:, sl]
    want_num = flat(ak.num(sel))[0]
    want_sum = flat(ak.sum(sel.pt, axis=-1))[0]
    want_arg = flat(ak.argmax(sel.pt, axis=-1))
    want_max = flat(ak.max(sel.pt, axis=-1))
    out = EventPipeline(ak.from_offsets(off, fields, hip), TERMS, sl, "pt").run()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    assert same_bits(want_num, got["num"])
    assert same_bits(want_sum, got["sum"])
    assert same_bits(want_arg[0], got["argmax_mask"]) and same_bits(want_arg[1], got["argmax"])
    assert same_bits(want_max[0], got["max_mask"]) and same_bits(want_max[1], got["max"])


def test_event_pipeline_long_events(cpu, hip):
    """Events whose particles overflow the LDS stage take the global path."""
    from awkward_amd.fused import EventPipeline

    off, fields = events(6, 2000, 40, 3000)
    ev_c = ak.from_offsets(off, fields, cpu)
    sel = ev_c[(ev_c.pt > 25) & (abs(ev_c.eta) < 2.4)][:, 1:]
    out = EventPipeline(ak.from_offsets(off, fields, hip), TERMS, slice(1, None), "pt").run()
    assert same_bits(flat(ak.num(sel))[0], out["num"].cpu().numpy())
    assert same_bits(flat(ak.sum(sel.pt, axis=-1))[0], out["sum"].cpu().numpy())


@pytest.mark.parametrize("lam,clip", [(4, 30), (14, 60), (0.7, 3)])
def test_pair_max_matches_elementwise_form(hip, lam, clip):
    """pair_max == ak.max/argmax over the materialised pairs with the same
    device expression (bit-exact), and the torch-ufunc form (tolerance).
    Poisson(4): the pairs of a workgroup fit its LDS tile (pair-parallel);
    Poisson(14): they do not (per-event enumeration); Poisson(0.7): most
    events have no pairs."""
    import torch

    from awkward_amd.fused import PairMax, pair_expr

    off, fields = events(7, 20000, lam, clip)
    ev = ak.from_offsets(off, fields, hip)
    pairs = ak.combinations(ev, 2)
    p0, p1 = pairs["0"], pairs["1"]
    legs0 = [p0[f].layout.content.project().data for f in ("pt", "eta", "phi")]
    legs1 = [p1[f].layout.content.project().data for f in ("pt", "eta", "phi")]
    expr = pair_expr(legs0, legs1)
    arr = ak.Array(ak.highlevel.C.ListOffsetArray(pairs.layout.offsets, ak.highlevel.C.NumpyArray(expr, hip), hip))
    m = flat(ak.max(arr, axis=-1))
    a = flat(ak.argmax(arr, axis=-1))
    out = PairMax(ev).run()
    torch.cuda.synchronize()
    assert same_bits(m[0], out["mask"].cpu().numpy()) and same_bits(m[1], out["max"].cpu().numpy())
    assert same_bits(a[1], out["argmax"].cpu().numpy())
    ufunc = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))
    u = ufunc.layout.content.data.cpu().numpy()
    assert np.allclose(u, expr.cpu().numpy(), rtol=2e-6, atol=0)


def test_concurrent_branches_in_a_graph(hip):
    """pair_max and the triples' combinations as two graph branches on their
    own streams: every replay equals the sequential eager results."""
    import torch

    from awkward_amd.fused import Combinations, Concurrent, Graph, PairMax

    off, fields = events(12, 40000, 4, 30)
    ev = ak.from_offsets(off, fields, hip)
    pm = PairMax(ev)
    tri = Combinations(ev, 3)
    ref_pm = {k: v.clone() for k, v in pm.run().items()}
    ref_tri = [c.clone() for c in tri.run()]
    torch.cuda.synchronize()
    g = Graph(Concurrent(tri.run, pm.run))
    for _ in range(3):
        for v in list(pm.out.values()) + tri.carry:
            v.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in ref_pm:
            assert torch.equal(pm.out[k], ref_pm[k])
        for a, b in zip(tri.carry, ref_tri):
            assert torch.equal(a, b)


def test_graph_replays_match_eager(hip):
    import torch

    from awkward_amd.fused import EventPipeline, Graph, PairMax

    off, fields = events(8, 50000, 3, 20)
    ev = ak.from_offsets(off, fields, hip)
    pipe = EventPipeline(ev, TERMS, slice(1, None), "pt")
    pm = PairMax(ev)

    def step():
        return pipe.run(), pm.run()

    eager = [{k: v.clone() for k, v in d.items()} for d in step()]
    g = Graph(step)
    for _ in range(4):
        for d in pipe.out, pm.out:
            for v in d.values():
                v.fill_(0)
        torch.cuda.synchronize()
        res = g.replay()
        torch.cuda.synchronize()
        for want, got in zip(eager, res):
            for k in want:
                assert same_bits(want[k].cpu().numpy(), got[k].cpu().numpy()), k


@pytest.mark.parametrize("n", [2, 3])
def test_combinations_runner_matches_registry_path(hip, n):
    from awkward_amd.fused import Combinations, Graph

    off, fields = events(9, 20000, 4, 30)
    ev = ak.from_offsets(off, fields, hip)
    want = ak.combinations(ev, n)
    runner = Combinations(ev, n)
    got = runner.run()
    wl = want.layout
    assert same_bits(wl.offsets.cpu().numpy(), runner.tooffsets.cpu().numpy())
    for k in range(n):
        idx = wl.content.contents[k].index
        assert same_bits(idx.cpu().numpy(), got[k][: runner.total].cpu().numpy())
    g = Graph(runner.run)
    for _ in range(3):
        for c in runner.carry:
            c.fill_(-1)
        g.replay()
        for k in range(n):
            assert same_bits(wl.content.contents[k].index.cpu().numpy(), runner.carry[k][: runner.total].cpu().numpy())


TERM_SHAPES = [
    [("pt", ">=", 25.0), ("eta", "abs<=", 2.4)],
    [("pt", "<", 40.0), ("phi", "!=", 0.5), ("eta", "abs>", 0.5)],
    [("eta", "abs<", 3.0), ("phi", "==", 0.5)],
    [],
    [("charge", ">", 0), ("pt", ">", 20.0)],  # an int8 column: the generic kernel
    [("pt", ">", 10.0), ("pt", "<", 200.0), ("eta", "abs<", 2.4), ("phi", ">", -3.0), ("phi", "<", 3.0)],
]


@pytest.mark.parametrize("sl", [slice(1, None), slice(None, None, 2)])
@pytest.mark.parametrize("k", range(len(TERM_SHAPES)))
def test_event_pipeline_term_shapes(cpu, hip, k, sl):
    """Every comparison, through abs or not, with NaN, infinities, signed
    zeros and exact hits of the constants in the columns; float32 cuts of at
    most four terms take the streamlined kernel, an int8 term or a fifth term
    the generic one; both walks ([start:] in one pass, a stepped slice in
    two) -- against the oracle backend's reference sequence."""
    from awkward_amd.fused import EventPipeline

    off, fields = events(20 + k, 20000, 3, 20)
    rng = np.random.default_rng(k)
    N = len(fields["pt"])
    for name in ("pt", "eta", "phi"):
        idx = rng.integers(0, N, N // 30)
        fields[name][idx] = rng.choice(np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 0.5, -0.5], np.float32),
                                       len(idx))
    terms = TERM_SHAPES[k]
    ev_c = ak.from_offsets(off, fields, cpu)
    mask = None
    for f, op, v in terms:
        x = abs(ev_c[f]) if op.startswith("abs") else ev_c[f]
        op = op[3:] if op.startswith("abs") else op
        m = {">": x > v, ">=": x >= v, "<": x < v, "<=": x <= v, "==": x == v, "!=": x != v}[op]
        mask = m if mask is None else mask & m
    sel = (ev_c[mask] if mask is not None else ev_c)[:, sl]
    out = EventPipeline(ak.from_offsets(off, fields, hip), terms, sl, "pt").run()
    got = {key: v.cpu().numpy() for key, v in out.items()}
    assert same_bits(flat(ak.num(sel))[0], got["num"])
    assert same_bits(flat(ak.sum(sel.pt, axis=-1))[0], got["sum"])
    want_arg = flat(ak.argmax(sel.pt, axis=-1))
    assert same_bits(want_arg[0], got["argmax_mask"]) and same_bits(want_arg[1], got["argmax"])
    want_max = flat(ak.max(sel.pt, axis=-1))
    assert same_bits(want_max[0], got["max_mask"]) and same_bits(want_max[1], got["max"])
