"""Batching form planning (host logic, no GPU): the C4 batch-size
distribution, the greedy pack planner and the serving step-size choice."""
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


def test_pick_step_chunks():
    from irismpc.batching import pick_step_chunks

    # seconds per chunk: larger steps amortise fixed costs, then flatten
    per = {1: 1.0, 2: 0.8, 4: 0.7, 8: 0.69, 16: 0.685}
    assert pick_step_chunks(per, tol=0.03) == 4  # 0.7 <= 0.685 * 1.03
    assert pick_step_chunks(per, tol=0.0) == 16
    assert pick_step_chunks(per, tol=0.5) == 1
    assert pick_step_chunks({1: 2.0}) == 1
    # a slower large step is never chosen over the fastest smaller one
    assert pick_step_chunks({1: 1.0, 2: 0.9, 4: 0.95}, tol=0.0) == 2
    # a step-time cap: k = 16 (11 s) and 8 (5.5 s) exceed 5 s, 4 is the best left
    assert pick_step_chunks(per, tol=0.03, max_step=5.0) == 4
    assert pick_step_chunks(per, tol=0.0, max_step=6.0) == 8
    assert pick_step_chunks(per, max_step=0.5) == 1  # nothing fits: the smallest step
