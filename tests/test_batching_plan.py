"""Batching form planning (host logic, no GPU): the C4 batch-size
distribution and the greedy pack planner."""
import numpy as np


def test_plan_packs_rules():
    from irismpc.batching import log_uniform_batch_sizes, plan_packs

    sizes = log_uniform_batch_sizes(200, seed=4)
    assert sizes.min() >= 1 and sizes.max() <= 256
    packs = plan_packs(sizes, 256, 64)
    assert [i for p in packs for i in p] == list(range(200))
    for p in packs:
        assert len(p) <= 64 and (len(p) == 1 or sum(sizes[i] for i in p) <= 256)


def test_counter_advance_matches_chunking():
    from irismpc.batching import counter_advance

    # two full chunks and a partial one, 62 rows: 2 * ceil(w * 62 * 2 / 64) each
    assert counter_advance(70000, 32768, 62) == 2 * (2 * -(-(32768 * 124) // 64)) + 2 * -(-(4464 * 124) // 64)
    assert counter_advance(0, 32768, 62) == 0


def test_range_advance_pads_the_last_chunk():
    from irismpc.batching import counter_advance, range_advance

    # the scan launches 32768, 32768 and 64-padded 4480 rows
    assert range_advance(70000, 32768, 62) == 2 * counter_advance(32768, 32768, 62) + counter_advance(4480, 32768, 62)
