"""Per-stream runtime state: scratch pools and error words belong to the
(device, stream) that calls the library, as the reference CUDA backend keeps
its error contexts per stream (src/awkward/_connect/cuda/__init__.py:28,
257-294).

* look-back scans enqueued on two streams at once, interleaved, each
  checked against the oracle (they would share tile counters and status
  words if the scratch were per device);
* a deferred-mode error recorded on one stream is reported by that stream's
  poll and not by the other's;
* a hipGraph captured on a stream keeps working after a larger eager call on
  the same stream grows that stream's pool (the old buffer is retired, not
  freed);
* the deferred-mode invocation table wraps without aliasing tags.
"""

import ctypes

import numpy as np
import pytest

from parity import offsets_from_counts

pytestmark = pytest.mark.gpu


def oracle_run(name, args):
    import oracle

    host = [a.copy() if isinstance(a, np.ndarray) else a for a in args]
    err = oracle.kernel(name)(*[a.ctypes.data if isinstance(a, np.ndarray) else a for a in host])
    return host, err


def compact_case(rng, n):
    counts = rng.poisson(5, n).astype(np.int64)
    starts = rng.integers(0, 1000, n).astype(np.int64)
    return starts, starts + counts


def test_two_streams_interleaved_scans():
    import torch

    from awkward_amd._kernels import HipKernel

    rng = np.random.default_rng(21)
    compact = HipKernel("awkward_ListArray64_compact_offsets_64")
    rng_len = HipKernel("awkward_ListArray64_getitem_next_range_64")
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    na, nb = 3_000_000, 2_000_000
    starts_a, stops_a = compact_case(rng, na)
    off_b = offsets_from_counts(rng.poisson(4, nb).astype(np.int64))
    da = [torch.from_numpy(x).cuda() for x in (starts_a, stops_a)]
    db = torch.from_numpy(off_b).cuda()
    total_b = int(np.maximum(np.diff(off_b) - 1, 0).sum())
    outs_a = [torch.empty(na + 1, dtype=torch.int64, device="cuda") for _ in range(4)]
    outs_b = [(torch.empty(nb + 1, dtype=torch.int64, device="cuda"),
               torch.empty(max(total_b, 1), dtype=torch.int64, device="cuda")) for _ in range(4)]
    torch.cuda.synchronize()
    from awkward_amd._lib import lib

    L = lib()
    L.awkward_hip_set_deferred_errors(1)  # no synchronisation per call: the scans stay in flight together
    try:
        for k in range(4):  # alternate launches: the two streams' scans overlap on the device
            with torch.cuda.stream(sa):
                assert compact(outs_a[k], da[0], da[1], na).str is None
            with torch.cuda.stream(sb):
                assert rng_len(outs_b[k][0], outs_b[k][1], db[:-1], db[1:], nb, 1, 2**63 - 1, 1).str is None
        torch.cuda.synchronize()
    finally:
        L.awkward_hip_set_deferred_errors(0)
    want_a, _ = oracle_run("awkward_ListArray64_compact_offsets_64", [np.zeros(na + 1, np.int64), starts_a,
                                                                     stops_a, na])
    want_b, _ = oracle_run("awkward_ListArray64_getitem_next_range_64",
                           [np.zeros(nb + 1, np.int64), np.zeros(max(total_b, 1), np.int64),
                            np.ascontiguousarray(off_b[:-1]), np.ascontiguousarray(off_b[1:]), nb, 1, 2**63 - 1, 1])
    for k in range(4):
        assert (outs_a[k].cpu().numpy() == want_a[0]).all(), k
        assert (outs_b[k][0].cpu().numpy() == want_b[0]).all(), k
        assert (outs_b[k][1].cpu().numpy()[:total_b] == want_b[1][:total_b]).all(), k


def test_deferred_errors_stay_on_their_stream():
    import torch

    from awkward_amd._kernels import HipKernel
    from awkward_amd._lib import lib

    L = lib()
    from awkward_amd._kernels import KernelError

    L.awkward_hip_error_poll.restype = KernelError
    compact = HipKernel("awkward_ListArray64_compact_offsets_64")
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    starts = torch.tensor([0, 5, 9], dtype=torch.int64, device="cuda")
    good_stops = torch.tensor([3, 7, 9], dtype=torch.int64, device="cuda")
    bad_stops = torch.tensor([3, 4, 9], dtype=torch.int64, device="cuda")  # stops[1] < starts[1]
    out_a = torch.empty(4, dtype=torch.int64, device="cuda")
    out_b = torch.empty(4, dtype=torch.int64, device="cuda")
    L.awkward_hip_set_deferred_errors(1)
    try:
        with torch.cuda.stream(sb):
            assert compact(out_b, starts, bad_stops, 3).str is None  # deferred: no error yet
        with torch.cuda.stream(sa):
            assert compact(out_a, starts, good_stops, 3).str is None
            from awkward_amd._kernels import set_stream

            set_stream()
            err_a = L.awkward_hip_error_poll()
        with torch.cuda.stream(sb):
            set_stream()
            err_b = L.awkward_hip_error_poll()
            err_b2 = L.awkward_hip_error_poll()
    finally:
        L.awkward_hip_set_deferred_errors(0)
    assert err_a.str is None
    assert err_b.str == b"stops[i] < starts[i]" and err_b.id == 1
    assert err_b2.str is None  # cleared by the first poll
    assert out_a.cpu().tolist() == [0, 3, 5, 5]


def test_graph_survives_pool_growth():
    import torch

    from awkward_amd._kernels import HipKernel
    from awkward_amd._lib import lib
    from awkward_amd.fused import Graph

    rng = np.random.default_rng(22)
    # next_range: a look-back scan with no host synchronisation, so it can be captured
    rng_len = HipKernel("awkward_ListArray64_getitem_next_range_64")
    n = 100_000
    off = offsets_from_counts(rng.poisson(4, n).astype(np.int64))
    do = torch.from_numpy(off).cuda()
    total = int(np.maximum(np.diff(off) - 1, 0).sum())
    out = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    carry = torch.empty(total, dtype=torch.int64, device="cuda")
    g = Graph(lambda: rng_len(out, carry, do[:-1], do[1:], n, 1, 2**63 - 1, 1))
    want, _ = oracle_run("awkward_ListArray64_getitem_next_range_64",
                         [np.zeros(n + 1, np.int64), np.zeros(total, np.int64), np.ascontiguousarray(off[:-1]),
                          np.ascontiguousarray(off[1:]), n, 1, 2**63 - 1, 1])
    L = lib()
    L.awkward_hip_pool_bytes.argtypes = [ctypes.c_void_p]
    L.awkward_hip_pool_bytes.restype = None
    pool = (ctypes.c_int64 * 3)()
    with torch.cuda.stream(g.stream):
        from awkward_amd._kernels import set_stream

        set_stream()
        L.awkward_hip_pool_bytes(pool)
        before = list(pool)
        # a sort of one long segment on the graph's stream needs far more
        # scratch (its merge buffers) than the scan: that stream's pool grows
        big = 4_000_000
        x = rng.integers(-10**9, 10**9, big).astype(np.int64)
        dx = torch.from_numpy(x).cuda()
        doff = torch.tensor([0, big], dtype=torch.int64, device="cuda")
        big_out = torch.empty(big, dtype=torch.int64, device="cuda")
        argsort = HipKernel("awkward_argsort_int64_int64")
        assert argsort(big_out, dx, big, doff, 2, True, True).str is None
        set_stream()
        L.awkward_hip_pool_bytes(pool)
        after = list(pool)
    torch.cuda.synchronize()
    assert after[0] > before[0] and after[2] > before[2], (before, after)
    for _ in range(3):
        out.fill_(-1)
        carry.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        assert (out.cpu().numpy() == want[0]).all()
        assert (carry.cpu().numpy() == want[1]).all()
    assert (big_out.cpu().numpy() == np.argsort(x, kind="stable")).all()


def test_deferred_invocation_table_wraps():
    """More than 65,536 deferred invocations without a poll: the error of an
    early call is still reported with its own message, and later calls'
    tags do not alias it."""
    import torch

    from awkward_amd._kernels import HipKernel, KernelError, set_stream
    from awkward_amd._lib import lib

    L = lib()
    L.awkward_hip_error_poll.restype = KernelError
    compact = HipKernel("awkward_ListArray64_compact_offsets_64")
    nextcarry = HipKernel("awkward_IndexedArray64_getitem_nextcarry_64")
    starts = torch.tensor([0, 5], dtype=torch.int64, device="cuda")
    bad_stops = torch.tensor([3, 4], dtype=torch.int64, device="cuda")
    out = torch.empty(3, dtype=torch.int64, device="cuda")
    index = torch.tensor([0, 1], dtype=torch.int64, device="cuda")
    carry = torch.empty(2, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    L.awkward_hip_set_deferred_errors(1)
    try:
        with torch.cuda.stream(s):
            assert compact(out, starts, bad_stops, 2).str is None
            for _ in range(70_000):
                assert nextcarry(carry, index, 2, 10).str is None
            set_stream()
            err = L.awkward_hip_error_poll()
    finally:
        L.awkward_hip_set_deferred_errors(0)
    assert err.str == b"stops[i] < starts[i]" and err.id == 1
