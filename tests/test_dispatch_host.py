"""Host-side tile choice (awkward_amd.dispatch): the long-list queue for
mean lengths of 2048 and more, ahead of any installed chooser; the fixed
heuristic's stage model otherwise; and the shape of the variant table the
library compiles (csrc/reduce.h, launch_reduce)."""

import pytest

from awkward_amd import dispatch


def test_long_lists_take_the_queue_before_any_chooser():
    calls = []
    saved = dispatch.chooser
    try:
        dispatch.chooser = lambda *a: calls.append(a) or 3
        assert dispatch.choose_variant("sum", 30, 30 * 2048, 8) == dispatch.LONG_LISTS
        assert dispatch.choose_variant("sum", 30, 30 * 2047, 8) == 3
        assert len(calls) == 1
    finally:
        dispatch.chooser = saved


@pytest.mark.parametrize("mean", [1, 4, 10, 40, 300, 2000])
def test_heuristic_stage_fits(mean):
    n = 100_000
    v = dispatch.heuristic_variant(n, n * mean, 8)
    assert 0 <= v < len(dispatch.VARIANTS)
    fill, _ = dispatch.variant_model(v, n, n * mean, 8)
    # the chosen stage is expected to hold its span, unless none does (then the largest stage)
    assert fill <= 0.85 or v == 4


def test_variant_table():
    assert len(dispatch.VARIANTS) == 10 and dispatch.LONG_LISTS == 10
    for block, lpt, cap in dispatch.VARIANTS:
        assert block % 64 == 0 and lpt >= 1 and cap % 1024 == 0
    # the 1M-list Poisson(10) float64 layout: a one-wave tile whose stage holds the span
    assert dispatch.heuristic_variant(1_000_000, 10_000_000, 8) == 6
