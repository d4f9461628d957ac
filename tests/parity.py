"""Runs one kernel specialization on the oracle (host buffers) and on
libawkward_hip.so (device buffers) with the same seeded inputs."""

from __future__ import annotations

import ctypes

import numpy as np

import oracle


def _host_arg(a, keep):
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, list):  # list of arrays -> table of pointers
        table = (ctypes.c_void_p * len(a))(*[x.ctypes.data for x in a])
        keep.append(table)
        return ctypes.addressof(table)
    return a


def _dev_arg(a, keep):
    import torch

    from awkward_amd._kernels import PointerArray

    if isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        keep.append(t)
        return t.data_ptr(), t
    if isinstance(a, list):
        ts = [torch.from_numpy(np.ascontiguousarray(x)).to("cuda") for x in a]
        table = PointerArray(ts)
        keep.append(table)
        return table.data_ptr(), ts
    return a, None


def run_both(name, args):
    """Returns (host_args_after, device_args_after, host_error, device_error);
    arrays are copied, so the caller's inputs are untouched."""
    import torch

    from awkward_amd._kernels import HipKernel

    host = [a.copy() if isinstance(a, np.ndarray) else ([x.copy() for x in a] if isinstance(a, list) else a)
            for a in args]
    keep = []
    herr = oracle.kernel(name)(*[_host_arg(a, keep) for a in host])
    dkeep, dcall, dobjs = [], [], []
    for a in args:
        ptr, obj = _dev_arg(a, dkeep)
        dcall.append(ptr)
        dobjs.append(obj)
    derr = HipKernel(name)._impl(*dcall)
    torch.cuda.synchronize()
    dev = []
    for a, obj in zip(args, dobjs):
        if isinstance(a, np.ndarray):
            dev.append(obj.cpu().numpy())
        elif isinstance(a, list):
            dev.append([t.cpu().numpy() for t in obj])
        else:
            dev.append(a)
    return host, dev, herr, derr


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        w = np.uint32 if a.dtype.itemsize == 4 else np.uint64
        return bool((a.view(w) == b.view(w)).all())
    return bool((a == b).all())


def offsets_from_counts(counts):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off


def mixed_counts(rng, n, long_every=0, long_len=5000, medium_every=0, medium_len=300, lam=6):
    """Poisson list lengths with some medium and long lists and empties."""
    counts = rng.poisson(lam, n).astype(np.int64)
    counts[rng.random(n) < 0.15] = 0
    if medium_every:
        counts[::medium_every] = rng.integers(33, medium_len, size=len(counts[::medium_every]))
    if long_every:
        counts[::long_every] = rng.integers(4097, long_len, size=len(counts[::long_every]))
    return counts
