"""The lazy front end: graphs recorded by LazyArray, evaluated eagerly
through the host layer or with the reductions that match the event pipeline
fused into one launch -- with the same bits either way."""

import numpy as np
import pytest

import awkward_amd as ak
from awkward_amd import lazy as L


def events(n, seed=2):
    import bench

    return bench.make_events(n, seed=seed)


def oracle_backend():
    from oracle.backend import OracleBackend

    return OracleBackend()


def same(a, b):
    """Same layout buffers, bit for bit (to_list plus the raw buffers)."""
    if a.to_list() != b.to_list():
        return False
    la, lb = a.layout, b.layout
    na, nb = la.nplike, lb.nplike
    if hasattr(la, "mask"):
        if not np.array_equal(na.to_numpy(la.mask).astype(np.int8), nb.to_numpy(lb.mask).astype(np.int8)):
            return False
        la, lb = la.content, lb.content
    x, y = na.to_numpy(la.data), nb.to_numpy(lb.data)
    return x.dtype == y.dtype and x.tobytes() == y.tobytes()


def pipeline_outputs(ev, sl=slice(1, None)):
    sel = ev[(ev.pt > 25) & (abs(ev.eta) < 2.4)][:, sl]
    return [sel.num(), sel.pt.sum(), sel.pt.argmax(), sel.pt.max(), sel.pt.count()]


def test_graph_records_the_operations():
    off, fields = events(100)
    ev = L.lazy(ak.from_offsets(off, fields, oracle_backend()))
    out = ev[(ev.pt > 25) & (abs(ev.eta) < 2.4)][:, 1:].pt.sum()
    text = out.visualize()
    for word in ("reduce", "field", "slice", "filter", "binary", "compare", "unary", "constant", "input"):
        assert word in text
    with pytest.raises(TypeError):
        bool(ev.pt > 1)


def test_eager_matches_the_host_layer():
    """fuse=False (and any backend without fused kernels) evaluates node by
    node through the same ak.* calls."""
    off, fields = events(5000)
    arr = ak.from_offsets(off, fields, oracle_backend())
    ev = L.lazy(arr)
    got = L.compute(*pipeline_outputs(ev))
    sel = arr[(arr.pt > 25) & (abs(arr.eta) < 2.4)][:, 1:]
    want = [ak.num(sel), ak.sum(sel.pt, axis=-1), ak.argmax(sel.pt, axis=-1), ak.max(sel.pt, axis=-1),
            ak.count(sel.pt, axis=-1)]
    for g, w in zip(got, want):
        assert same(g, w)
    assert L.Plan(pipeline_outputs(ev)).stats()["fused_launches"] == 0  # not a hip array
    combos = ev.combinations(2).compute()
    assert combos.to_list() == ak.combinations(arr, 2).to_list()


@pytest.mark.gpu
@pytest.mark.parametrize("sl", [slice(1, None), slice(None), slice(0, 2), slice(None, None, 2)])
def test_fused_equals_eager(sl):
    from awkward_amd.backend import HipBackend

    off, fields = events(50000)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    outs = pipeline_outputs(ev, sl)
    plan = L.Plan(outs)
    stats = plan.stats()
    assert stats["fused_reductions"] == 5 and stats["fused_launches"] == 1 and stats["eager_outputs"] == 0
    fused = plan.run()
    eager = L.compute(*outs, fuse=False)
    for f, e in zip(fused, eager):
        assert same(f, e)


@pytest.mark.gpu
def test_mixed_plan_and_graph_replay():
    from awkward_amd.backend import HipBackend

    off, fields = events(30000, seed=5)
    ev = L.lazy(ak.from_offsets(off, fields, HipBackend()))
    sel = ev[(ev.pt > 30) & (ev.charge > 0)]
    outs = [sel.pt.sum(), sel.pt.min(), sel.num()]  # min is not a fused output: eager
    plan = L.Plan(outs)
    assert plan.stats()["fused_reductions"] == 2 and plan.stats()["eager_outputs"] == 1
    got = plan.run()
    want = L.compute(*outs, fuse=False)
    for g, w in zip(got, want):
        assert same(g, w)
    fused_only = L.Plan([sel.pt.sum(), sel.pt.argmax()])
    graph = fused_only.graph()
    first = [v.clone() for v in graph.replay()[0].values()]
    for _ in range(3):
        again = graph.replay()[0]
        assert all(bool((a == b).all()) for a, b in zip(first, again.values()))
    with pytest.raises(ValueError):
        plan.graph()
