"""Batch processing end to end (ServerActor::process_batch_query,
iris-mpc-gpu/src/server/actor.rs:674-1790): two batches against a 1,000-row
DB of both eyes. The expected outcome of every request is written out, and
for the first batch the DB matches and the batch self-deduplication are also
derived from the oracle's plaintext brute force (both eyes at 0.345, the
OR rule, the deleted row) and must agree with both. Modelled on the request kinds of the reference's e2e generator
(iris-mpc-common/src/test.rs): unique insertions, DB matches, batch
duplicates, skip_persistence, successful and failed reauths, a left-only
match under the AND rule and under the OR rule, a deletion, and in the second
batch queries of irises written by the first (insertions, a reauth) and of the
deleted row.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N0, MAXB, MAXDB = 1000, 16, 4096
NM = 0xFFFFFFFF


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


def _iris(ds, i):
    from irismpc import synth

    return synth.synth_iris(ds, i)


def _flip(code, key, n=1280):
    from irismpc import synth

    return synth.synth_flip(code, 900, key, n)


class Batch:
    def __init__(self):
        self.ids, self.types, self.irises = [], [], {"left": [], "right": []}
        self.skip, self.reauth, self.or_rule = [], {}, []

    def add(self, rid, left, right, kind="uniqueness", skip=False, reauth=None, or_rule=()):
        self.ids.append(rid)
        self.types.append(kind)
        self.irises["left"].append(left)
        self.irises["right"].append(right)
        self.skip.append(skip)
        self.or_rule.append(list(or_rule))
        if reauth is not None:
            self.reauth[rid] = reauth

    def query(self, deletions=()):
        from irismpc import synth
        from irismpc.server import BatchQuery

        sh = {}
        for k, eye in enumerate(("left", "right")):
            codes = np.array([c for c, _ in self.irises[eye]])
            masks = np.array([m for _, m in self.irises[eye]])
            sh[eye] = synth.shares_of(codes, masks, 500 + k + 2 * len(self.ids))
        return BatchQuery(self.ids, self.types, sh["left"], sh["right"], self.or_rule, self.skip,
                          self.reauth, list(deletions))


def test_server_two_batches(irismpc, orc):
    from irismpc import synth
    from irismpc.server import ServerActor

    db = {"left": [_iris(11, i) for i in range(N0)], "right": [_iris(12, i) for i in range(N0)]}
    actor = ServerActor(MAXDB, MAXB)
    sh = {}
    for k, eye in enumerate(("left", "right")):
        codes = np.array([c for c, _ in db[eye]])
        masks = np.array([m for _, m in db[eye]])
        sh[eye] = synth.shares_of(codes, masks, 300 + k)
    actor.load_db(sh["left"], sh["right"])
    assert actor.current_db_size == N0

    def copy(eye, i, key):
        c, m = db[eye][i]
        return _flip(c, key), m

    new = {n: (_iris(21, n), _iris(22, n)) for n in range(6)}  # fresh irises (left, right)
    fresh = lambda n, key: ((_flip(new[n][0][0], key), new[n][0][1]), (_flip(new[n][1][0], key), new[n][1][1]))  # noqa: E731

    b1 = Batch()
    b1.add("u0", new[0][0], new[0][1])                                     # 0 unique -> row 1000
    b1.add("m17", copy("left", 17, 1), copy("right", 17, 2))               # 1 DB match 17
    b1.add("dup0", *fresh(0, 3))                                            # 2 duplicate of 0
    b1.add("skip1", new[1][0], new[1][1], skip=True)                       # 3 unique, not stored
    b1.add("re33", copy("left", 33, 4), copy("right", 33, 5), kind="reauth", reauth=33)  # 4 ok
    b1.add("re40", new[2][0], new[2][1], kind="reauth", reauth=40)         # 5 fails
    b1.add("l50", copy("left", 50, 6), _iris(23, 0))                       # 6 left only -> unique, row 1001
    b1.add("or60", copy("left", 60, 7), _iris(23, 1), or_rule=[60])        # 7 OR rule -> 60
    b1.add("u3", new[3][0], new[3][1])                                     # 8 unique -> row 1002
    b1.add("dup3", *fresh(3, 8))                                            # 9 duplicate of 8
    b1.add("d70", copy("left", 70, 9), copy("right", 70, 10))              # 10 deleted row -> unique, 1003
    b1.add("u4", new[4][0], new[4][1])                                     # 11 unique -> row 1004
    r1 = actor.process_batch_query(b1.query(deletions=[70]))

    FB = 0xFFFFFFFE
    assert r1.merged_results == [1000, 17, FB, NM, 33, NM, 1001, 60, 1002, FB, 1003, 1004]
    assert r1.matches == [False, True, True, True, True, True, False, True, False, True, False, False]
    assert r1.matches_with_skip_persistence[3] is False and r1.matches_with_skip_persistence[1] is True
    assert r1.matched_batch_request_ids[2] == ["u0"] and r1.matched_batch_request_ids[9] == ["u3"]
    assert all(not r1.matched_batch_request_ids[i] for i in (0, 1, 3, 4, 5, 6, 7, 8, 10, 11))
    assert r1.successful_reauths == [i == 4 for i in range(12)]
    assert r1.match_ids[1] == [17] and r1.match_ids[4] == [33] and r1.match_ids[7] == [60]
    assert r1.match_ids[2] == [] and r1.match_ids[9] == []  # batch matches are reported by request id
    assert r1.partial_match_ids_left[6] == [50] and r1.partial_match_ids_right[6] == []
    assert r1.partial_match_ids_left[7] == [60] and r1.partial_match_counters_left[7] == 1
    assert r1.partial_match_ids_left[10] == [] and r1.partial_match_ids_right[10] == []  # deleted
    assert r1.partial_match_rotation_indices_left[1] == [[0]]
    assert r1.deleted_ids == [70]
    assert actor.current_db_size == N0 + 5

    # the same answers from the plaintext brute force (integer MPC predicate, 0.345)
    A = irismpc.THRESHOLD_A
    plain_db = {e: (np.array([c for c, _ in db[e]]), np.array([m for _, m in db[e]])) for e in ("left", "right")}
    for q in range(len(b1.ids)):
        hit = {e: orc.plain_match(*b1.irises[e][q], *plain_db[e], A) for e in ("left", "right")}
        rows = set(np.flatnonzero(hit["left"] & hit["right"]).tolist())
        rows |= {d for d in b1.or_rule[q] if hit["left"][d] or hit["right"][d]}
        assert set(r1.match_ids[q]) == rows - {70}, f"request {b1.ids[q]}: DB matches"
        dup = [b1.ids[j] for j in range(q)
               if all(orc.plain_match(*b1.irises[e][q], *(np.array([x]) for x in b1.irises[e][j]), A)[0]
                      for e in ("left", "right"))]
        assert r1.matched_batch_request_ids[q] == dup, f"request {b1.ids[q]}: batch duplicates"

    b2 = Batch()
    b2.add("q_u0", *fresh(0, 20))                       # 0 -> row 1000 (inserted in batch 1)
    b2.add("q_skip1", *fresh(1, 21))                    # 1 -> not stored before: unique, row 1005
    b2.add("q_re33", copy("left", 33, 22), copy("right", 33, 23))  # 2 -> 33
    b2.add("q_d70", copy("left", 70, 24), copy("right", 70, 25))   # 3 -> 1003, no longer 70
    b2.add("q_u3", *fresh(3, 26))                       # 4 -> 1002
    b2.add("q_l50", copy("left", 50, 27), _iris(23, 0))  # 5 -> 1001 (its left and right)
    b2.add("q_u4", *fresh(4, 28), kind="reauth", reauth=1004)      # 6 reauth of an inserted row
    r2 = actor.process_batch_query(b2.query())
    assert r2.merged_results == [1000, 1005, 33, 1003, 1002, 1001, 1004]
    assert r2.successful_reauths == [False] * 6 + [True]
    assert r2.partial_match_ids_left[3] == [1003] and r2.partial_match_ids_right[3] == [1003]
    assert actor.current_db_size == N0 + 6


def test_server_first_batch_on_empty_db(irismpc):
    """The reference's server starts from an empty DB: its first batch has no
    DB matches, inserts its unique irises and flags in-batch duplicates; the
    next batch finds the inserted rows."""
    from irismpc.server import ServerActor

    actor = ServerActor(MAXDB, MAXB)
    assert actor.current_db_size == 0
    new = {n: (_iris(31, n), _iris(32, n)) for n in range(2)}
    b1 = Batch()
    b1.add("u0", new[0][0], new[0][1])
    b1.add("dup0", (_flip(new[0][0][0], 40), new[0][0][1]), (_flip(new[0][1][0], 41), new[0][1][1]))
    b1.add("u1", new[1][0], new[1][1])
    r1 = actor.process_batch_query(b1.query())
    FB = 0xFFFFFFFE
    assert r1.merged_results == [0, FB, 1]
    assert r1.matches == [False, True, False]
    assert r1.matched_batch_request_ids[1] == ["u0"]
    assert actor.current_db_size == 2
    b2 = Batch()
    b2.add("q1", (_flip(new[1][0][0], 42), new[1][0][1]), (_flip(new[1][1][0], 43), new[1][1][1]))
    r2 = actor.process_batch_query(b2.query())
    assert r2.merged_results == [1] and r2.matches == [True]
    assert actor.current_db_size == 2


def test_server_mirror_attack(irismpc, orc):
    """Full-face mirror attack detection (actor.rs:585-640,1304-1417): a face
    presented mirrored shows each eye as the mirror image of the other. The
    Mirror pass (each eye's DB against the other eye's query mirrored) finds
    it, the Normal pass does not; it is flagged and not inserted."""
    from irismpc import synth
    from irismpc.server import ServerActor

    N = 300
    db = {"left": [_iris(31, i) for i in range(N)], "right": [_iris(32, i) for i in range(N)]}
    actor = ServerActor(1024, 8)
    sh = {}
    for k, eye in enumerate(("left", "right")):
        sh[eye] = synth.shares_of(np.array([c for c, _ in db[eye]]), np.array([m for _, m in db[eye]]), 310 + k)
    actor.load_db(sh["left"], sh["right"])

    def mirrored(eye, i, key):
        c, m = orc.iris_mirrored(*db[eye][i])
        return _flip(c, key, 800), m

    def build():
        b = Batch()
        b.add("attack", mirrored("right", 10, 1), mirrored("left", 10, 2))      # mirror image of face 10
        b.add("m20", (_flip(db["left"][20][0], 3), db["left"][20][1]),
              (_flip(db["right"][20][0], 4), db["right"][20][1]))               # ordinary DB match
        b.add("new", _iris(33, 0), _iris(34, 0))                                # unique
        return b

    q = build().query()
    q.full_face_mirror_attacks_detection_enabled = True
    r = actor.process_job(q)
    assert r.merged_results == [NM, 20, N]
    assert r.full_face_mirror_attack_detected == [True, False, False]
    assert r.full_face_mirror_match_ids[0] == [10] and r.full_face_mirror_match_ids[1] == []
    assert r.matches == [True, True, False]
    assert actor.current_db_size == N + 1

    # without detection the mirrored face is taken as a new identity
    q = build().query()
    r = actor.process_job(q)
    assert r.merged_results[0] == N + 1 and r.full_face_mirror_attack_detected == [False] * 3


def test_server_anon_stats(irismpc):
    """Anonymized-statistics distances of a batch (actor.rs:1631-1650): the
    two-sided cache holds the (batch, DB row, query) pairs captured in both
    eyes: 0.375 matches and reauth target pairs, matching or not; a left-only
    match stays one-sided. All parties hold the same keys."""
    from irismpc import synth
    from irismpc.server import ServerActor

    N, MAXDB, B = 300, 1024, 8
    db = {"left": [_iris(51, i) for i in range(N)], "right": [_iris(52, i) for i in range(N)]}
    actor = ServerActor(MAXDB, B, match_distances_buffer_size=1 << 14)
    sh = {}
    for k, eye in enumerate(("left", "right")):
        sh[eye] = synth.shares_of(np.array([c for c, _ in db[eye]]), np.array([m for _, m in db[eye]]), 320 + k)
    actor.load_db(sh["left"], sh["right"])
    copy = lambda eye, i, key: (_flip(db[eye][i][0], key), db[eye][i][1])  # noqa: E731
    b = Batch()
    b.add("m10", copy("left", 10, 1), copy("right", 10, 2))
    b.add("l20", copy("left", 20, 3), _iris(53, 0))
    b.add("re30", copy("left", 30, 4), copy("right", 30, 5), kind="reauth", reauth=30)
    b.add("re40", _iris(53, 1), _iris(54, 1), kind="reauth", reauth=40)
    r = actor.process_batch_query(b.query())
    # query 1 matches in one eye only: unique, inserted as row N
    assert r.merged_results[:3] == [10, N, 30] and r.successful_reauths == [False, False, True, False]
    key = lambda d, j: actor.internal_batch_counter * MAXDB * B + d * B + j  # noqa: E731
    for p in range(3):
        assert set(r.two_sided_match_distances[p].map) == {key(10, 0), key(30, 2), key(40, 3)}
        assert key(20, 1) in r.one_sided_match_distances["left"][p].map
        left, right = r.two_sided_match_distances[p].map[key(40, 3)]
        assert len(left) == len(right) == 31  # every rotation of the reauth pair


def test_server_checks_luc_and_identity_updates(irismpc, orc):
    """Reset and recovery checks are matched like uniqueness requests but never
    inserted; LUC look-back (luc_lookback_records = 20) adds the last DB rows
    to the OR rule of uniqueness requests only, so a left-eye-only copy of row
    990 matches while the same kind of reset check does not; an identity update
    replaces row 40 before the batch is matched; a batch of updates only
    (a deletion and an identity update) changes the DB without matching; every
    DB match is also checked against the plaintext brute force."""
    from irismpc import synth
    from irismpc.server import ServerActor

    db = {"left": [_iris(31, i) for i in range(N0)], "right": [_iris(32, i) for i in range(N0)]}
    actor = ServerActor(MAXDB, MAXB)
    sh = {}
    for k, eye in enumerate(("left", "right")):
        sh[eye] = synth.shares_of(np.array([c for c, _ in db[eye]]), np.array([m for _, m in db[eye]]), 310 + k)
    actor.load_db(sh["left"], sh["right"])

    def copy(eye, i, key, src=None):
        c, m = (src or db[eye])[i]
        return _flip(c, key), m

    def both_shares(left, right, key):
        out = []
        for k, (c, m) in enumerate((left, right)):
            s = synth.shares_of(c[None], m[None], 700 + 2 * key + k)
            out.append([(s[p][0][0], s[p][1][0]) for p in range(3)])
        return tuple(out)

    x = (_iris(33, 0), _iris(34, 0))  # the new iris of row 40 (left, right)
    b = Batch()
    b.add("rc17", copy("left", 17, 1), copy("right", 17, 2), kind="reset_check")
    b.add("rcnew", _iris(35, 0), _iris(36, 0), kind="reset_check")
    b.add("rec33", copy("left", 33, 3), copy("right", 33, 4), kind="recovery_check")
    b.add("unew", _iris(35, 1), _iris(36, 1))
    b.add("l990", copy("left", 990, 5), _iris(37, 0))
    b.add("rcl995", copy("left", 995, 6), _iris(37, 1), kind="reset_check")
    b.add("qx", (_flip(x[0][0], 7), x[0][1]), (_flip(x[1][0], 8), x[1][1]))
    b.add("qold40", copy("left", 40, 9), copy("right", 40, 10))
    q = b.query()
    q.luc_lookback_records = 20
    q.identity_update_request_ids = ["upd40"]
    q.identity_update_indices = [40]
    q.identity_update_shares = [both_shares(x[0], x[1], 0)]
    r = actor.process_batch_query(q)

    FB = 0xFFFFFFFE
    assert r.merged_results == [17, NM, 33, 1000, 990, NM, 40, 1001]
    assert r.matches == [True, True, True, False, True, True, True, False]
    assert actor.current_db_size == N0 + 2
    assert r.identity_update_indices == [40] and r.identity_update_request_ids == ["upd40"]
    assert FB not in r.merged_results
    # brute force over the DB as matched (row 40 replaced), OR rule = LUC rows for uniqueness requests
    A = irismpc.THRESHOLD_A
    plain = {e: [list(v) for v in db[e]] for e in ("left", "right")}
    plain["left"][40], plain["right"][40] = list(x[0]), list(x[1])
    plain = {e: (np.array([c for c, _ in v]), np.array([m for _, m in v])) for e, v in plain.items()}
    luc = set(range(N0 - 1 - 20, N0))
    for j, kind in enumerate(b.types):
        hit = {e: orc.plain_match(*b.irises[e][j], *plain[e], A) for e in ("left", "right")}
        rows = set(np.flatnonzero(hit["left"] & hit["right"]).tolist())
        if kind == "uniqueness":
            rows |= {d for d in luc if hit["left"][d] or hit["right"][d]}
        assert set(r.match_ids[j]) == rows, f"{b.ids[j]}: DB matches"

    # a batch of non-MPC updates only
    y = (_iris(33, 1), _iris(34, 1))
    from irismpc.server import BatchQuery

    empty = [(np.zeros((0, 12800), np.uint16), np.zeros((0, 6400), np.uint16))] * 3
    u = BatchQuery([], [], empty, empty, deletion_requests_indices=[17],
                   identity_update_request_ids=["upd33"], identity_update_indices=[33],
                   identity_update_shares=[both_shares(y[0], y[1], 1)])
    ru = actor.process_batch_query(u)
    assert ru.deleted_ids == [17] and ru.identity_update_indices == [33] and ru.merged_results == []
    assert actor.current_db_size == N0 + 2
    c = Batch()
    c.add("q17", copy("left", 17, 11), copy("right", 17, 12))
    c.add("qy", (_flip(y[0][0], 13), y[0][1]), (_flip(y[1][0], 14), y[1][1]))
    rc = actor.process_batch_query(c.query())
    assert rc.merged_results == [N0 + 2, 33]


def test_server_refuses_disagreeing_parties(irismpc):
    """The three parties' merged results are compared before any result is
    assembled or written (the role of sync_match_results, actor.rs:2591-2630):
    a party whose first match differs makes the batch fail and the DB stays
    as it was."""
    from irismpc import synth
    from irismpc.server import ServerActor

    db = {"left": [_iris(41, i) for i in range(64)], "right": [_iris(42, i) for i in range(64)]}
    actor = ServerActor(1024, 4)
    sh = {}
    for k, eye in enumerate(("left", "right")):
        sh[eye] = synth.shares_of(np.array([c for c, _ in db[eye]]), np.array([m for _, m in db[eye]]), 410 + k)
    actor.load_db(sh["left"], sh["right"])
    b = Batch()
    b.add("u", _iris(43, 0), _iris(44, 0))
    b.add("m5", (_flip(db["left"][5][0], 1), db["left"][5][1]), (_flip(db["right"][5][0], 2), db["right"][5][1]))
    match = actor.matcher.match

    def tampered(*a, **k):
        r = match(*a, **k)
        r.final_results[1] = r.final_results[1].copy()
        r.final_results[1][1] ^= 1
        return r

    actor.matcher.match = tampered
    with pytest.raises(RuntimeError, match="disagree"):
        actor.process_batch_query(b.query())
    assert actor.current_db_size == 64  # nothing inserted
    actor.matcher.match = match
    r = actor.process_batch_query(b.query())
    assert r.merged_results == [64, 5]
