"""Wave planning of the batch runner (awkward_amd/batch.py) on the CPU:
consecutive arrays grouped under a byte budget, every array in exactly one
wave, an oversized array alone, and the per-array byte model."""

import numpy as np

from awkward_amd.batch import device_bytes, plan_waves


def test_waves_cover_in_order_within_budget():
    rng = np.random.default_rng(0)
    sizes = list(rng.integers(1, 100, 200))
    budget = 250
    waves = plan_waves(sizes, budget)
    assert [i for w in waves for i in w] == list(range(len(sizes)))
    for w in waves:
        assert sum(sizes[i] for i in w) <= budget or len(w) == 1
    # greedy: the next array would not have fit in the previous wave
    for a, b in zip(waves, waves[1:]):
        assert sum(sizes[i] for i in a) + sizes[b[0]] > budget


def test_oversized_array_is_a_wave_of_its_own():
    assert plan_waves([5, 5, 50, 5], 20) == [[0, 1], [2], [3]]
    assert plan_waves([], 10) == []


def test_device_bytes_model():
    # offsets, int32 content, sum + min + argmin + mask per list, argsort out + sort scratch per element
    n, N = 1000, 50_000
    assert device_bytes(n, N, 4, argsort=False) == 8 * (n + 1) + 4 * N + n * (8 + 4 + 8 + 1)
    assert device_bytes(n, N, 4) == device_bytes(n, N, 4, argsort=False) + 32 * N
