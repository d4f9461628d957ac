"""Batch runner (persistent work-queue kernels) against the oracle, per array."""

import numpy as np
import pytest

from parity import offsets_from_counts, run_both, same_bits

pytestmark = pytest.mark.gpu


def make_batch(seed, narrays=24, max_exp=4.3):
    """A small batch with the batched workload's mix: list counts log-uniform, list
    lengths Poisson(2) / Poisson(50) / Zipf(1.5) capped, some all-empty
    arrays, int32 / int64 / float32 / float64 values."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(narrays):
        n = int(10 ** rng.uniform(1, max_exp))
        kind = rng.integers(0, 3)
        if kind == 0:
            counts = rng.poisson(2, n)
        elif kind == 1:
            counts = rng.poisson(50, n)
        else:
            counts = np.minimum(rng.zipf(1.5, n), 6000)
        if k % 7 == 3:
            counts = np.zeros(n, dtype=np.int64)
        off = offsets_from_counts(counts)
        dt = [np.int32, np.int64, np.float32, np.float64][k % 4]
        N = int(off[-1])
        if np.dtype(dt).kind == "i":
            x = rng.integers(-1000, 1000, N).astype(dt)
        else:
            x = rng.standard_normal(N).astype(dt)
        out.append((off, x))
    return out


def check_array(off, x, res):
    n = len(off) - 1
    t = x.dtype.name
    sum_t = {"int32": "int64", "int64": "int64", "float32": "float32", "float64": "float64"}[t]
    h, _, _, _ = run_both(f"awkward_reduce_sum_{sum_t}_{t}_64", [np.zeros(n, sum_t), x, off, n])
    got = res["sum"].cpu().numpy()
    if x.dtype.kind == "i":
        assert same_bits(h[0], got)
    else:
        counts = np.diff(off)
        eps = np.finfo(x.dtype).eps
        absum = np.array([np.abs(x[a:b].astype(np.float64)).sum() for a, b in zip(off[:-1], off[1:])])
        assert (np.abs(h[0].astype(np.float64) - got.astype(np.float64)) <= 2 * counts * eps * absum).all()
    ident = int(np.iinfo(x.dtype).max) if x.dtype.kind == "i" else np.inf
    h, _, _, _ = run_both(f"awkward_reduce_min_{t}_{t}_64", [np.zeros(n, t), x, off, n, ident])
    assert same_bits(h[0], res["min"].cpu().numpy())
    h, _, _, _ = run_both(f"awkward_reduce_argmin_{t}_64", [np.zeros(n, np.int64), x, off, off[:-1].copy(), n])
    want = np.where(h[0] >= 0, h[0] - off[:-1], h[0])
    assert same_bits(want, res["argmin"].cpu().numpy())
    assert same_bits((np.diff(off) == 0).astype(np.int8), res["mask"].cpu().numpy())
    h, _, _, _ = run_both(f"awkward_argsort_{t}_int64", [np.zeros(len(x), np.int64), x, len(x), off, n + 1, True, True])
    assert same_bits(h[0], res["argsort"].cpu().numpy())


@pytest.mark.parametrize("variant", range(10))
def test_batch_runner_matches_oracle(variant):
    import torch

    from awkward_amd.batch import BatchRunner

    batch = make_batch(40 + variant)
    dev = [(torch.from_numpy(o).cuda(), torch.from_numpy(x).cuda()) for o, x in batch]
    runner = BatchRunner(dev, variants={k: variant for k in ("int32", "int64", "float32", "float64")})
    runner.run()
    torch.cuda.synchronize()
    for i, (off, x) in enumerate(batch):
        check_array(off, x, runner.result(i))


def test_batch_graph_replays():
    import torch

    from awkward_amd.batch import BatchRunner
    from awkward_amd.fused import Graph

    batch = make_batch(50)
    dev = [(torch.from_numpy(o).cuda(), torch.from_numpy(x).cuda()) for o, x in batch]
    runner = BatchRunner(dev)
    runner.run()
    torch.cuda.synchronize()
    eager = [{k: v.clone() for k, v in runner.result(i).items()} for i in range(len(batch))]
    g = Graph(runner.run)
    for _ in range(3):
        for grp in runner.groups:
            for out in grp.outputs.values():
                for v in out.values():
                    v.fill_(7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(len(batch)):
            got = runner.result(i)
            for k, v in eager[i].items():
                assert torch.equal(v, got[k]), (i, k)
    timings = runner.tune(reps=1)
    assert set(timings) <= {"int32", "int64", "float32", "float64"}


def test_batch_long_lists():
    """Lists of 5e3..1e5 elements: the queued long-segment sort across
    arrays, and the reductions' long-list path."""
    import torch

    from awkward_amd.batch import BatchRunner

    rng = np.random.default_rng(60)
    batch = []
    for k, dt in enumerate([np.float32, np.int64, np.float64, np.int32, np.float32]):
        counts = rng.integers(4000, 100001, 6 + k)
        counts[0] = 4097
        off = offsets_from_counts(counts)
        N = int(off[-1])
        x = rng.integers(-1000, 1000, N).astype(dt) if np.dtype(dt).kind == "i" else rng.standard_normal(N).astype(dt)
        batch.append((off, x))
    dev = [(torch.from_numpy(o).cuda(), torch.from_numpy(x).cuda()) for o, x in batch]
    runner = BatchRunner(dev)
    runner.run()
    torch.cuda.synchronize()
    for i, (o, x) in enumerate(batch):
        check_array(o, x, runner.result(i))


def test_waves_straddle_the_budget():
    """A batch run in waves (a budget that splits it into several runs of
    arrays, one array alone because it exceeds the budget): every array's
    results equal the oracle's, as in one run."""
    import torch

    from awkward_amd.batch import BatchRunner, WaveRunner, device_bytes

    batch = make_batch(77, narrays=14)
    sizes = [device_bytes(len(o) - 1, len(x), x.itemsize) for o, x in batch]
    budget = sorted(sizes)[len(sizes) // 2] * 2
    specs = [(len(o) - 1, len(x), x.dtype) for o, x in batch]
    waves = WaveRunner(specs, lambda i: (torch.from_numpy(batch[i][0]).cuda(), torch.from_numpy(batch[i][1]).cuda()),
                       budget)
    assert len(waves.waves) >= 3 and sorted(i for w in waves.waves for i in w) == list(range(len(batch)))
    assert all(waves.bytes(w) <= budget or len(w) == 1 for w in waves.waves)
    seen = 0
    for wave, arrays in waves:
        runner = BatchRunner(arrays)
        runner.run()
        torch.cuda.synchronize()
        for k, i in enumerate(wave):
            check_array(batch[i][0], batch[i][1], runner.result(k))
            seen += 1
    assert seen == len(batch)
