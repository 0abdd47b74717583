"""Host side of the anonymized-statistics distance caches
(iris-mpc-gpu/src/server/anon_stats.rs:136-403): ordering, filtering and
grouping of captured entries, and the two-sided merge. Cases built by hand
from the reference's definitions."""
import numpy as np

from irismpc.anon_stats import (REAUTH, UNIQUENESS, OneSidedDistanceCache, TwoSidedDistanceCache,
                                load_additions, operation_from_match_id, target_idx_from_match_id)

QMAX, DBMAX = 4 * 31, 100  # max batch 4


class _Fake:
    def __init__(self, idx):
        self.idx = np.array(idx, np.uint64)

    def fetch(self):
        n = len(self.idx)
        v = np.arange(n, dtype=np.uint16)
        return n, self.idx, v, v + 100, v + 200, v + 300


def _mid(batch, db, row):
    return batch * DBMAX * QMAX + db * QMAX + row


def test_match_id_helpers():
    i = _mid(3, 17, 2 * 31 + 5)
    assert operation_from_match_id(i, [UNIQUENESS, UNIQUENESS, REAUTH], QMAX) == REAUTH
    assert operation_from_match_id(i, [], QMAX) == UNIQUENESS
    assert target_idx_from_match_id(i, QMAX, DBMAX) == 17 * 4 + 2


def test_load_additions_filters_and_groups():
    ids = [_mid(1, 9, 31 + 3), _mid(1, 9, 31 + 2),   # query 1, db 9: two rotations
           _mid(1, 4, 0),                              # query 0, db 4
           _mid(1, 5, 2 * 31),                          # query 2 (skip_persistence)
           _mid(1, 6, 3 * 31 + 1), _mid(1, 7, 3 * 31)]  # query 3 reauth: target db 7 only
    ops = [UNIQUENESS, UNIQUENESS, UNIQUENESS, REAUTH]
    skip = [False, False, True, False]
    c = load_additions(_Fake(ids), 0, 1000, ops, skip, [7 * 4 + 3], QMAX, DBMAX)
    keys = sorted(c.map)
    assert keys == sorted({_mid(1, 9, 31) // 31, _mid(1, 4, 0) // 31, _mid(1, 7, 93) // 31})
    group = c.map[_mid(1, 9, 31) // 31]
    assert [e.idx for e in group] == sorted([ids[0], ids[1]])  # sorted by match id
    e = group[0]
    assert (e.code_a, e.code_b, e.mask_a, e.mask_b) == (1, 101, 201, 301)  # entry 1 of the buffer
    assert c.map[_mid(1, 7, 93) // 31][0].operation == REAUTH
    # a buffer filled past the limit yields nothing; old entries are skipped
    assert len(load_additions(_Fake(ids), 0, 6, ops, skip, [], QMAX, DBMAX)) == 0
    assert len(load_additions(_Fake(ids[:2]), 1, 1000, ops, skip, [], QMAX, DBMAX).map) == 1


def test_two_sided_merge_and_extend():
    left = OneSidedDistanceCache({1: ["l1"], 2: ["l2"], 5: ["l5"]})
    right = OneSidedDistanceCache({2: ["r2"], 5: ["r5"], 9: ["r9"]})
    two = TwoSidedDistanceCache.merge(left, right)
    assert two.map == {2: (["l2"], ["r2"]), 5: (["l5"], ["r5"])}
    acc = TwoSidedDistanceCache()
    acc.extend(two)
    assert len(acc) == 2
    try:
        acc.extend(two)
        raise AssertionError("duplicate keys must be refused")
    except AssertionError as err:
        assert "already exist" in str(err)
