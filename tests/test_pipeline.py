"""End-to-end pipelines on the hip backend against the same host layer
running on the oracle backend (CPU), at sizes the oracle finishes quickly."""

import numpy as np
import pytest

import awkward_amd as ak
from parity import same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["fused", "plain"])
def backends(request):
    """The oracle backend and a hip backend: fused (one tuned kernel per
    reduction) or plain (the reference's kernel sequence)."""
    from awkward_amd.backend import HipBackend
    from oracle.backend import OracleBackend

    return OracleBackend(), HipBackend(fused=request.param == "fused")


def events(seed, n, lam, clip):
    rng = np.random.default_rng(seed)
    mult = np.clip(rng.poisson(lam, n), 0, clip).astype(np.int64)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(mult, out=off[1:])
    N = int(off[-1])
    fields = {
        "pt": (5 + rng.exponential(20, N)).astype(np.float32),
        "eta": rng.uniform(-2.5, 2.5, N).astype(np.float32),
        "phi": rng.uniform(-np.pi, np.pi, N).astype(np.float32),
        "charge": (2 * rng.integers(0, 2, N) - 1).astype(np.int8),
    }
    return off, fields


def flat(array):
    """Raw buffers of a result, for bit-exact comparison."""
    layout = array.layout
    nplike = layout.nplike
    from awkward_amd import contents as C

    if isinstance(layout, C.ByteMaskedArray):
        return [nplike.to_numpy(layout.mask), nplike.to_numpy(layout.content.data)]
    if isinstance(layout, C.NumpyArray):
        return [nplike.to_numpy(layout.data)]
    if isinstance(layout, C.ListOffsetArray):
        inner = layout.content
        if isinstance(inner, C.IndexedArray):
            inner = inner.project()
        return [nplike.to_numpy(layout.offsets), nplike.to_numpy(inner.data)]
    raise TypeError(layout)


def event_pipeline_steps(ev):
    mask = (ev.pt > 25) & (abs(ev.eta) < 2.4)
    sel = ev[mask][:, 1:]
    return {
        "num": ak.num(sel),
        "sum": ak.sum(sel.pt, axis=-1),
        "argmax": ak.argmax(sel.pt, axis=-1),
        "pt": sel.pt,
    }


def test_event_pipeline(backends):
    cpu, hip = backends
    off, fields = events(2, 20000, 3, 20)
    out_c = event_pipeline_steps(ak.from_offsets(off, fields, cpu))
    out_h = event_pipeline_steps(ak.from_offsets(off, fields, hip))
    for k in out_c:
        for a, b in zip(flat(out_c[k]), flat(out_h[k])):
            assert same_bits(a, b), k


def pair_pipeline_steps(ev, n):
    pairs = ak.combinations(ev, n)
    p0, p1 = pairs["0"], pairs["1"]
    expr = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))
    return expr, ak.max(expr, axis=-1), ak.argmax(expr, axis=-1)


@pytest.mark.parametrize("n", [2, 3])
def test_pair_pipeline(backends, n):
    cpu, hip = backends
    off, fields = events(3, 5000, 4, 30)
    e_c, m_c, a_c = pair_pipeline_steps(ak.from_offsets(off, fields, cpu), n)
    e_h, m_h, a_h = pair_pipeline_steps(ak.from_offsets(off, fields, hip), n)
    oc, vc = flat(e_c)
    oh, vh = flat(e_h)
    assert same_bits(oc, oh)
    # cosh/cos come from different math libraries: tolerance on the values,
    # and the reductions are checked on the device's own expression values
    assert np.allclose(vc, vh, rtol=1e-5, atol=1e-3)
    ev_h = ak.Array(ak.highlevel.C.ListOffsetArray(oh, ak.highlevel.C.NumpyArray(vh, cpu), cpu))
    for a, b in zip(flat(ak.max(ev_h, axis=-1)), flat(m_h)):
        assert same_bits(a, b)
    for a, b in zip(flat(ak.argmax(ev_h, axis=-1)), flat(a_h)):
        assert same_bits(a, b)


def test_segmented_reductions_1m_layout(backends):
    cpu, hip = backends
    rng = np.random.default_rng(1)
    counts = rng.poisson(10, 20000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    x = rng.standard_normal(int(off[-1]))
    a_c = ak.from_offsets(off, x, cpu)
    a_h = ak.from_offsets(off, x, hip)
    for fn in (ak.sum, ak.max, ak.argmax, ak.count, ak.min, ak.argmin):
        for a, b in zip(flat(fn(a_c, axis=-1)), flat(fn(a_h, axis=-1))):
            assert same_bits(a, b), fn.__name__


@pytest.mark.parametrize("ascending", [True, False])
def test_argsort_sort_api(backends, ascending):
    cpu, hip = backends
    rng = np.random.default_rng(4)
    counts = rng.poisson(50, 2000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    x = rng.integers(-1000, 1000, int(off[-1])).astype(np.int32)
    for fn in (ak.argsort, ak.sort):
        c = fn(ak.from_offsets(off, x, cpu), ascending=ascending)
        h = fn(ak.from_offsets(off, x, hip), ascending=ascending)
        for a, b in zip(flat(c), flat(h)):
            assert same_bits(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["poisson", "long_lists", "offset_start", "unaligned_mask", "all_false", "all_true",
                                  "empty_lists"])
def test_bool_to_index_kernel(case):
    """awkward_hip_bool_to_index (tiles of 3072 lists whose mask bytes are
    staged in LDS, or read from global memory when a tile's span exceeds the
    12 KiB stage) against numpy: the local positions of the true values per
    list and their offsets (the reference's _slicing.py:476-530 steps)."""
    import torch

    from awkward_amd.nplike import TorchHip

    rng = np.random.default_rng(77)
    n = 50_000
    counts = rng.poisson(3, n)
    if case == "long_lists":
        sel = rng.random(n) < 0.3
        counts[sel] = rng.integers(15, 60, int(sel.sum()))
        counts[::997] = 5000
    if case == "empty_lists":
        counts[rng.random(n) < 0.7] = 0
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    total = int(offsets[-1])
    mask = rng.random(total + 7) < 0.4
    if case == "all_false":
        mask[:] = False
    if case == "all_true":
        mask[:] = True
    if case == "offset_start":
        offsets = offsets + 5  # the lists start 5 bytes into the mask
    big = torch.from_numpy(mask).cuda()
    d_mask = big[3:] if case == "unaligned_mask" else big
    host_mask = mask[3:] if case == "unaligned_mask" else mask
    d_off = torch.from_numpy(offsets).cuda()
    nl = TorchHip()
    index, tooffsets = nl.bool_to_jagged_index(d_off, d_mask)
    want_index, want_off = [], [0]
    for a, b in zip(offsets[:-1], offsets[1:]):
        sel = np.nonzero(host_mask[a:b])[0]
        want_index.append(sel)
        want_off.append(want_off[-1] + len(sel))
    want_index = np.concatenate(want_index) if want_index else np.zeros(0, np.int64)
    assert np.array_equal(tooffsets.cpu().numpy(), np.array(want_off, dtype=np.int64)), case
    assert np.array_equal(index.cpu().numpy(), want_index.astype(np.int64)), case
