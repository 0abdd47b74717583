"""Host result assembly of the batch loop (iris-mpc-gpu/src/server/actor.rs:
3174-3303): merged results across devices, spreading insertions over
devices, new ids of inserted queries. Cases worked out by hand from the
reference's definitions, multi-device included."""
import numpy as np
import pytest

from irismpc.server import (NON_MATCH_ID, BatchQuery, calculate_insertion_indices,
                            distribute_insertions, get_merged_results)


def test_get_merged_results_devices():
    NM = NON_MATCH_ID
    # device-local ids -> global id = local * n_devices + device; min over devices
    host = [[5, NM, 0, NM], [2, NM, NM, 7], [NM, NM, NM, 1]]
    assert get_merged_results(host, 3) == [2 * 3 + 1, NM, 0, 1 * 3 + 2]
    assert get_merged_results([[NM, 4]], 1) == [NM, 4]
    # a batch match (0xFFFFFFFE) on one device stays itself
    assert get_merged_results([[0xFFFFFFFE]], 1) == [0xFFFFFFFE]


def test_distribute_insertions_starts_at_smallest_db():
    assert distribute_insertions([3, 5, 8, 9], [10, 7, 9]) == [[8], [3, 9], [5]]
    assert distribute_insertions([], [1, 1]) == [[], []]
    assert distribute_insertions([0, 1], [4]) == [[0, 1]]


def test_calculate_insertion_indices():
    merged = [NON_MATCH_ID] * 6
    ins = [[4], [1, 5], [2]]
    m = calculate_insertion_indices(merged, ins, [10, 7, 9], 6)
    # device order starts at the list whose first insertion is earliest (1 on
    # device 1): 1 -> 26, 2 -> 27, 4 -> 28, then the second sweep: 5 -> 29
    assert merged[1] == 26 and merged[2] == 27 and merged[4] == 28 and merged[5] == 29
    assert m == [True, False, False, True, False, False]
    one = [NON_MATCH_ID, 3, NON_MATCH_ID]
    assert calculate_insertion_indices(one, [[0, 2]], [100], 3) == [False, True, False]
    assert one == [100, 3, 101]


def test_batch_query_retain():
    import numpy as np

    sh = [(np.arange(6).reshape(3, 2), np.arange(3).reshape(3, 1)) for _ in range(3)]
    b = BatchQuery(["a", "b", "c"], ["uniqueness", "reauth", "uniqueness"], sh, sh,
                   [[1], [2], [3]], [False, True, False], {"b": 7}, [9])
    r = b.retain([0, 2])
    assert r.request_ids == ["a", "c"] and r.reauth_target_indices == {}
    assert r.or_rule_indices == [[1], [3]] and r.skip_persistence == [False, False]
    assert r.left[0][0].tolist() == [[0, 1], [4, 5]] and r.deletion_requests_indices == [9]


def test_or_policy_bitmap_known_answers():
    """iris-mpc-gpu/tests/or_policy.rs:8-33: rows of ceil(8000 / 64) = 125
    words; query 0's ids 1, 5, 10 -> word 0 = 1058, query 1 none -> word 125 = 0,
    query 2's id 3 -> word 250 = 8."""
    from irismpc.two_stage import prepare_or_policy_bitmap

    bm = prepare_or_policy_bitmap(8000, [[1, 5, 10], [], [3]], 3)
    assert bm.size * 64 == 8000 * 3
    assert int(bm[0]) == 1058 and format(int(bm[0]), "064b").endswith("10000100010")
    assert int(bm[125]) == 0 and int(bm[250]) == 8
    assert int(np.count_nonzero(bm)) == 2


@pytest.mark.parametrize("latest,or_rule,lookback,batch,skip,expected", [
    # test_only_lookback_records
    (10, [], 5, 2, set(), [[5, 6, 7, 8, 9, 10], [5, 6, 7, 8, 9, 10]]),
    # test_empty_zero_lookback_records
    (10, [[1, 3], [4, 5, 9]], 0, 1, set(), [[1, 3], [4, 5, 9]]),
    # test_basic_lookback
    (10, [[1, 3], [4, 5, 9]], 3, 2, set(), [[1, 3, 7, 8, 9, 10], [4, 5, 7, 8, 9, 10]]),
    # test_duplicate_ids
    (5, [[1, 1, 2, 3], [3, 3, 4]], 2, 2, set(), [[1, 2, 3, 4, 5], [3, 4, 5]]),
    # test_skip_lookback_requests
    (10, [[1, 2], [4, 5]], 3, 2, {0}, [[1, 2], [4, 5, 7, 8, 9, 10]]),
    # saturating start (latest < lookback)
    (2, [[7]], 5, 1, set(), [[0, 1, 2, 7]]),
])
def test_generate_luc_records_known_answers(latest, or_rule, lookback, batch, skip, expected):
    """generate_luc_records against iris-mpc-gpu/tests/or_policy.rs:36-147."""
    from irismpc.server import generate_luc_records

    assert generate_luc_records(latest, or_rule, lookback, batch, skip) == expected
