"""The no-allocation entry point
awkward_hip_ListArray64_getitem_next_range_64_noalloc against the entry
point it mirrors, awkward_ListArray64_getitem_next_range_64, through the
C ABI: bit-exact offsets and carries on two calls in a row, the second with
a smaller input and another slice (so anything the first call left in the
shared scratch array must have been cleared), then a third, larger again.
The size limit is checked on the host, without a GPU.
"""

import ctypes

import numpy as np
import pytest

KSLICENONE = 2**63 - 1
MAX_LISTS = ((1 << 20) - 512) // 8 * 512 * 8  # status words x the scan tile (512 threads x 8 lists)


def _fn(name):
    from awkward_amd._kernels import KernelError
    from awkward_amd._lib import lib

    f = getattr(lib(), name)
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] * 4
    f.restype = KernelError
    return f


def test_rejects_sizes_beyond_the_scratch():
    f = _fn("awkward_hip_ListArray64_getitem_next_range_64_noalloc")
    err = f(None, None, None, None, MAX_LISTS + 1, 1, KSLICENONE, 1)
    assert err.str is not None and b"exceeds the no-allocation scratch" in err.str


@pytest.mark.gpu
def test_matches_the_allocating_entry_point():
    import torch

    from awkward_amd._kernels import set_stream

    noalloc = _fn("awkward_hip_ListArray64_getitem_next_range_64_noalloc")
    plain = _fn("awkward_ListArray64_getitem_next_range_64")
    rng = np.random.default_rng(31)
    cases = [(3_000_000, (1, KSLICENONE, 1)), (70_000, (KSLICENONE, -1, 2)), (5_000_000, (-2, KSLICENONE, 1))]
    for n, (start, stop, step) in cases:
        counts = rng.poisson(4, n).astype(np.int64)
        counts[rng.random(n) < 0.1] = 0
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=off[1:])
        d_off = torch.from_numpy(off).cuda()
        starts, stops = d_off[:-1], d_off[1:]
        total = int(off[-1])  # a carry never exceeds the elements
        outs = []
        for f in (plain, noalloc):
            to = torch.full((n + 1,), -7, dtype=torch.int64, device="cuda")
            tc = torch.full((max(total, 1),), -7, dtype=torch.int64, device="cuda")
            set_stream()
            err = f(to.data_ptr(), tc.data_ptr(), starts.data_ptr(), stops.data_ptr(), n, start, stop, step)
            assert err.str is None, err.str
            torch.cuda.synchronize()
            outs.append((to.cpu().numpy(), tc.cpu().numpy()))
        (po, pc), (qo, qc) = outs
        assert (po == qo).all(), n
        k = int(po[-1])
        assert (pc[:k] == qc[:k]).all(), n
        assert (qc[k:] == -7).all(), n  # nothing written past the carry


@pytest.mark.gpu
def test_empty_input():
    import torch

    from awkward_amd._kernels import set_stream

    noalloc = _fn("awkward_hip_ListArray64_getitem_next_range_64_noalloc")
    to = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    set_stream()
    assert noalloc(to.data_ptr(), 0, 0, 0, 0, 1, KSLICENONE, 1).str is None
    torch.cuda.synchronize()
    assert to.item() == 0
