"""Serving form: continuous batching over the DB scan, with reauth, insertion
and deletion requests in the mix.

Checked on a 4-chunk DB of synthetic irises (4,096 distinct irises tiled):
- every uniqueness result in the mix equals the result of the same request
  alone (the plain per-request form) and the oracle's plaintext brute force
  over the tiled DB with the integer MPC predicate;
- reauth against the right / a wrong row; an inserted iris is found by a
  later query; a deleted row no longer matches;
- requests overlap in the active set; admission by free HBM lets one request
  run at a time when the reserve leaves no room;
- step sizes chosen by the autotuner leave every result unchanged;
- a server starting from an empty DB (as the reference's does) answers
  uniqueness requests with no matches, then finds an inserted iris;
- requests in flight together are compared with each other (the reference's
  batch self-comparison): each reports the earlier-admitted ones it matches,
  against the oracle's plaintext predicate between the request irises.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, U = 4 * 32768, 4096


@pytest.fixture(scope="module")
def served():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc
    from irismpc import serving, synth

    irismpc.lib()
    eng = serving.ServingEngines(N, N + 2 * 32768, max_active=8, seed_base=5)
    eng.build(U, 71, 72)
    return eng, serving, synth


def _query(synth, ds_iris, t, flip_idx, flips, rand_idx):
    c, m = synth.synth_iris(ds_iris, t)
    if flips:
        c = synth.synth_flip(c, 75, flip_idx, flips)
    sh = synth.shares_of(c[None], m[None], 76, rand_idx)
    return [(sh[p][0][0], sh[p][1][0]) for p in range(3)], c, m


def _brute_force(orc, c, m, deleted=(), extra=None):
    """Rows of the served DB (iris (71, r mod U) in row r < N, deleted rows
    excepted, `extra` {row: (code, mask)} after them) that the query matches
    at the 0.375 prefilter: the oracle's plaintext brute force."""
    global _DB_PLAIN
    if _DB_PLAIN is None:
        _DB_PLAIN = tuple(map(np.array, zip(*[orc.synth_iris(71, u) for u in range(U)])))
    hit = orc.plain_match(c, m, *_DB_PLAIN, 16384)
    rows = [r for r in np.flatnonzero(np.tile(hit, N // U)).tolist() if r not in deleted]
    for row, (ec, em) in sorted((extra or {}).items()):
        if orc.plain_match(c, m, ec[None], em[None], 16384)[0]:
            rows.append(row)
    return rows


_DB_PLAIN = None


def test_serving_mix_matches_alone(served, orc):
    eng, serving, synth = served
    R = serving.Request
    rng = np.random.default_rng(3)
    reqs, expect = [], {}
    targets = rng.choice(U, size=6, replace=False)
    deleted = int(targets[2]) + 2 * U  # one copy of iris targets[2], deleted at t = 0
    # uniqueness: 6 planted (10-30 % flips), 4 fresh irises
    for k in range(10):
        if k < 6:
            sh, c, m = _query(synth, 71, int(targets[k]), k, int(rng.integers(1280, 3841)), k)
        else:
            sh, c, m = _query(synth, 73, k, 0, 0, k)
        expect[k] = _brute_force(orc, c, m, deleted={deleted})
        reqs.append(R("uniqueness", 0.0, shares=sh, rid=k))  # one burst: they share the active set
    assert all(expect[k] for k in range(6)) and not any(expect[k] for k in range(6, 10))
    # reauth: right row, wrong row
    sh, _, _ = _query(synth, 71, int(targets[0]), 50, 1500, 50)
    reqs.append(R("reauth", 0.001, shares=sh, target=int(targets[0]) + U, rid=20))
    reqs.append(R("reauth", 0.003, shares=sh, target=int(targets[1]), rid=21))
    # insertion of a fresh iris, deletion of one copy of targets[2]
    new_c, new_m = synth.synth_iris(74, 1)
    ins = synth.shares_of(new_c[None], new_m[None], 77, 1)
    reqs.append(R("insert", 0.0, shares=[(ins[p][0][0], ins[p][1][0]) for p in range(3)], rid=30))
    reqs.append(R("delete", 0.0, target=deleted, rid=31))
    # after both applied: a query for the inserted iris, one for the deleted row's iris
    sh_new, c_new, m_new = _query(synth, 74, 1, 60, 2000, 60)
    sh_del, c_del, m_del = _query(synth, 71, int(targets[2]), 61, 2000, 61)
    reqs.append(R("uniqueness", 0.6, shares=sh_new, rid=40))
    reqs.append(R("uniqueness", 0.6, shares=sh_del, rid=41))

    srv = serving.Server(eng, reserve_bytes=0)
    out = {r.rid: r for r in srv.run(reqs)}
    assert srv.max_concurrent >= 4, "requests did not overlap in the active set"
    for k in range(10):
        assert out[k].result == expect[k], f"uniqueness {k} in the mix"
    assert out[20].result is True and out[21].result is False
    assert out[30].result == N and eng.n_rows == N + 1
    ins = {N: (new_c, new_m)}
    assert out[40].result == _brute_force(orc, c_new, m_new, {deleted}, ins) == [N]
    assert out[41].result == _brute_force(orc, c_del, m_del, {deleted}, ins)
    # the same requests alone (plain per-request form)
    for k in (0, 3, 7):
        assert eng.run_alone(reqs[k].shares) == out[k].result, f"uniqueness {k} alone"
    assert eng.run_alone(sh_new) == [N]


def test_admission_by_free_hbm(served):
    import torch

    eng, serving, synth = served
    R = serving.Request
    reqs = []
    for k in range(4):
        sh, _, _ = _query(synth, 73, 100 + k, 0, 0, 100 + k)
        reqs.append(R("uniqueness", 0.0, shares=sh, rid=k))
    free, _ = torch.cuda.mem_get_info()
    srv = serving.Server(eng, reserve_bytes=free + (1 << 30))  # no room beyond the first request
    out = srv.run(reqs)
    assert all(r.result == [] for r in out)
    assert srv.max_concurrent == 1 and srv.admission_waits > 0


def test_tuned_step_sizes_keep_results(served, tmp_path):
    eng, serving, synth = served
    from irismpc.batching import Autotuner

    sh, _, _ = _query(synth, 71, 5, 90, 2000, 90)
    plain = eng.run_alone(sh)
    assert len(plain) >= N // U - 1 and all(r % U == 5 for r in plain)  # one copy may be deleted
    table = eng.tune_steps(Autotuner(str(tmp_path / "tune.json")))
    try:
        assert set(table) == {1, 2, 4, 8}  # buckets up to max_active = 8
        for b, k in table.items():
            assert k >= 1 and k * b <= eng.max_active and k <= eng.n_chunks(eng.n_rows)
        assert eng.run_alone(sh) == plain
        # cached: a second tuner on the same file returns the same table without timing
        assert Autotuner(str(tmp_path / "tune.json")).step_chunks(eng) == table
    finally:
        eng.step_table = None


def test_empty_db_then_insert(served):
    _, serving, synth = served
    import irismpc

    eng = serving.ServingEngines(0, 2 * irismpc.DB_CHUNK_SIZE, max_active=2, seed_base=9)
    sh, _, _ = _query(synth, 78, 3, 0, 0, 3)
    assert eng.run_alone(sh) == []
    R = serving.Request
    out = serving.Server(eng, reserve_bytes=0).run([R("uniqueness", 0.0, shares=sh, rid=0),
                                                    R("insert", 0.0, shares=sh, rid=1)])
    res = {r.rid: r for r in out}
    assert res[1].result == 0 and eng.n_rows == 1
    assert res[0].result in ([], [0])  # checked before or after the insertion, as it was admitted
    near, _, _ = _query(synth, 78, 3, 5, 1500, 4)  # the inserted iris with 1,500 code bits flipped
    assert eng.run_alone(near) == [0]


def test_default_rids_do_not_alias(served, orc):
    """Two uniqueness requests left at the default rid, with different
    answers: the request loop tells them apart (it keys its active set on the
    requests themselves), so each is scanned with its own shares, served one
    after the other and together."""
    eng, serving, synth = served
    R = serving.Request
    sh_a, c_a, m_a = _query(synth, 71, 11, 70, 1500, 70)
    sh_b, c_b, m_b = _query(synth, 71, 12, 71, 1500, 71)
    ea, eb = _brute_force(orc, c_a, m_a), _brute_force(orc, c_b, m_b)
    assert ea and eb and not set(ea) & set(eb)
    alone_a, alone_b = eng.run_alone(sh_a), eng.run_alone(sh_b)
    # an earlier test deletes one DB row: at most one brute-force row is gone
    for got, exp in ((alone_a, ea), (alone_b, eb)):
        assert set(got) <= set(exp) and len(got) >= len(exp) - 1
    a, b = R("uniqueness", 0.0, shares=sh_a), R("uniqueness", 0.0, shares=sh_b)
    assert a.rid == b.rid == 0
    assert serving.Server(eng, reserve_bytes=0).run([a])[0].result == alone_a
    assert serving.Server(eng, reserve_bytes=0).run([b])[0].result == alone_b
    a2, b2 = R("uniqueness", 0.0, shares=sh_a), R("uniqueness", 0.01, shares=sh_b)
    out = serving.Server(eng, reserve_bytes=0).run([a2, b2])
    assert out[0].result == alone_a and out[1].result == alone_b


def test_short_bitmap_rows_rejected(served):
    """A scan step whose bitmap rows are shorter than the DB is refused with an
    error before anything is launched (the served DB, engines of its own)."""
    import torch

    eng, serving, synth = served
    from irismpc import engine as E

    sh, _, _ = _query(synth, 71, 3, 80, 1500, 80)
    ex = E.Exchange(timeout_s=60)
    base = [np.full(8, 40 + p, np.uint32) for p in range(3)]
    parties, qc, qm, streams = [], [], [], []
    for p in range(3):
        seeds = (E.derive_seed(base[p], 0), E.derive_seed(base[(p + 2) % 3], 0))
        codes = E.ShareDB(p, 32768, 31, E.IRIS_CODE_LENGTH, seeds, comm=ex, tag=0)
        masks = E.ShareDB(p, 32768, 31, E.MASK_CODE_LENGTH, seeds, comm=ex, tag=1)
        circ = E.Circuits(p, 31 * 32768, 31 * 32768 // 64, seeds, ex, tag=2)
        parties.append(E.PartyEngines(codes, masks, circ, E.DistanceComparator(31, eng.capacity),
                                      eng.code_dbs[p], eng.mask_dbs[p]))
        qc.append(codes.upload_query_shares(sh[p][0][None]))
        qm.append(masks.upload_query_shares(sh[p][1][None]))
        streams.append(E.stream_create())
    n = eng.n_rows
    bitmaps = [torch.zeros(n // 64 + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
    slots = [torch.zeros(1, dtype=torch.int32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="shorter than the DB"):
        E.scan_chunks_trio(parties, qc, qm, 0, 1, n, 1, eng.A, [b.data_ptr() for b in bitmaps], streams,
                           [m.data_ptr() for m in slots], row_bits=n - 64)
    for s in streams:
        E.stream_destroy(s)


def test_native_loop_rejects_bad_requests(served):
    """Requests the loop cannot serve fail with their reason (the run raises
    after every request finished); the loop itself keeps serving."""
    eng, serving, synth = served
    R = serving.Request
    sh, _, _ = _query(synth, 73, 200, 0, 0, 200)
    n = eng.n_rows
    bad = [R("reauth", 0.0, shares=sh, target=n + 5, rid=1), R("delete", 0.0, target=-3, rid=2)]
    with pytest.raises(RuntimeError, match="2 request\\(s\\) failed"):
        serving.Server(eng, reserve_bytes=0).run(bad)
    assert "outside the DB" in bad[0].error and "outside the DB" in bad[1].error
    ok = serving.Server(eng, reserve_bytes=0).run([R("uniqueness", 0.0, shares=sh, rid=3)])
    assert ok[0].error is None and ok[0].result == eng.run_alone(sh)


def test_per_batch_form_matches_serving_form(served):
    """The reference's per-batch form (batches padded to max_batch queries) on
    the same engines gives every request the result of the serving form."""
    eng, serving, synth = served
    R = serving.Request
    reqs = []
    for k in range(5):
        sh, _, _ = _query(synth, 71, 300 + k, 90 + k, 1500, 90 + k)
        reqs.append(R("uniqueness", 0.001 * k, shares=sh, rid=k))
    sh, _, _ = _query(synth, 71, 302, 99, 1500, 99)
    reqs.append(R("reauth", 0.002, shares=sh, target=302 + U, rid=10))
    srv = serving.PerBatchServer(eng, 4)
    out = {r.rid: r for r in srv.run([R(r.kind, r.arrival, shares=r.shares, target=r.target, rid=r.rid)
                                      for r in reqs])}
    assert srv.batches >= 2  # six requests, at most four per batch
    ref = {r.rid: r for r in serving.Server(eng, reserve_bytes=0).run(reqs)}
    for k in range(5):
        assert out[k].result == ref[k].result == eng.run_alone(reqs[k].shares), f"uniqueness {k}"
        assert out[k].result, "planted copies match"
    assert out[10].result is True and ref[10].result is True


def test_native_loop_records_and_release(served):
    """Through the C ABI directly: a request submitted ahead of its arrival is
    held until then; its result stays readable until released, and a pending
    request cannot be released."""
    import irismpc

    eng, serving, synth = served
    sh, _, _ = _query(synth, 73, 400, 0, 0, 400)
    eng.start(serving.CONTINUOUS, 0, 0)
    try:
        t0 = eng.now()
        i = eng.submit(serving.Request("uniqueness", 0.0, shares=sh), t0 + 0.3)
        with pytest.raises(irismpc.IrisMpcError, match="still pending"):
            serving.call("irismpc_serving_release", eng.h, i)
    finally:
        eng.finish(60)
    r = serving.Request("uniqueness", 0.0, shares=sh)
    eng.fetch(i, r, t0)
    assert r.error is None and r.result == [] and r.done >= 0.3  # held until its arrival
    serving.call("irismpc_serving_release", eng.h, i)
    with pytest.raises(irismpc.IrisMpcError, match="unknown request id"):
        eng.fetch(i, r, t0)


def _versioned_db_matches(orc, c, m, base, writes, version):
    """Brute force over the DB after the first `version` writes: base = (codes,
    masks, n_rows, U) of the tiled DB (row r holds distinct iris r mod U);
    writes = [(kind, row, code, mask)] in the loop's order. A deleted row holds
    the default iris (code 0, mask all ones) that deletions write."""
    codes, masks, n_rows, u = base
    hit = np.tile(orc.plain_match(c, m, codes, masks, 16384), -(-n_rows // u))[:n_rows].tolist()
    dummy = orc.plain_match(c, m, np.zeros((1, 200), np.uint64), np.full((1, 200), ~np.uint64(0)), 16384)[0]
    for kind, row, wc, wm in writes[:version]:
        if kind == "insert":
            assert row == len(hit)
            hit.append(bool(orc.plain_match(c, m, wc[None], wm[None], 16384)[0]))
        else:
            hit[row] = bool(dummy)
    return [r for r, h in enumerate(hit) if h]


@pytest.mark.parametrize("form", ["serving", "per_batch"])
def test_answers_follow_the_writes_in_order(orc, form):
    """Insertions and deletions land while uniqueness requests scan (one chunk
    per step, so every request spans several steps): a chunk written during a
    request's scan is scanned again for the whole active set, so each answer
    equals the oracle's brute force over the DB after exactly the request's
    db_version writes, replayed in the loop's order -- whichever writes
    happened mid-scan. The per-batch form answers for the DB before each
    batch. Rows: 3 chunks less one, so the second insertion opens a new
    chunk."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from irismpc import serving, synth

    n0, u = 3 * 32768 - 1, 1024
    eng = serving.ServingEngines(n0, n0 + 3 * 32768, max_active=8, seed_base=41)
    eng.build(u, 91, 92)
    eng.step_table = {1: 1}
    plain = tuple(map(np.array, zip(*[orc.synth_iris(91, i) for i in range(u)])))
    R = serving.Request
    rng = np.random.default_rng(17)
    # irises the writes change: X and Z inserted, two copies of Y's row class deleted
    xc, xm = synth.synth_iris(93, 0)
    zc, zm = synth.synth_iris(93, 1)
    y = int(rng.integers(u))

    def shares(c, m, k):
        sh = synth.shares_of(c[None], m[None], 94, k)
        return [(sh[p][0][0], sh[p][1][0]) for p in range(3)]

    reqs, queries, k = [], {}, 0
    for j in range(8):  # uniqueness every millisecond: X, Y, Z (10-25 % flips) in turn, two fresh irises
        which = j % 4
        if which == 3:
            qc, qm = synth.synth_iris(95, j)
        else:
            bc, bm = ((xc, xm), (plain[0][y], plain[1][y]), (zc, zm))[which]
            qc, qm = synth.synth_flip(bc, 96, j, int(rng.integers(1280, 3201))), bm
        queries[k] = (qc, qm)
        reqs.append(R("uniqueness", 0.001 * j, shares=shares(qc, qm, k), rid=k))
        k += 1
    writes_in = [("insert", 0.0005, xc, xm), ("delete", 0.0015, y + u, None), ("insert", 0.0025, zc, zm),
                 ("delete", 0.0035, y + 5 * u, None), ("insert", 0.0045, *synth.synth_iris(93, 2))]
    for kind, t, a, b in writes_in:
        if kind == "insert":
            reqs.append(R("insert", t, shares=shares(a, b, k), rid=k))
        else:
            reqs.append(R("delete", t, target=a, rid=k))
        k += 1
    srv = serving.Server(eng, reserve_bytes=0) if form == "serving" else serving.PerBatchServer(eng, 4)
    out = {r.rid: r for r in srv.run(reqs)}
    writes = sorted((r for r in out.values() if r.kind in ("insert", "delete")), key=lambda r: r.db_version)
    assert [w.db_version for w in writes] == [1, 2, 3, 4, 5]
    replay = []
    for w in writes:
        kind, _, a, b = writes_in[w.rid - 8]
        replay.append((kind, w.result, a, b) if kind == "insert" else (kind, w.result, None, None))
    assert [w.result for w in writes if w.kind == "insert"] == [n0, n0 + 1, n0 + 2]  # the second opens chunk 3
    base = (plain[0], plain[1], n0, u)
    mid_scan = 0
    for j in range(8):
        r = out[j]
        assert r.error is None
        assert r.result == _versioned_db_matches(orc, *queries[j], base, replay, r.db_version), \
            f"uniqueness {j} at DB version {r.db_version}"
        mid_scan += sum(1 for w in writes if w.arrival > r.arrival and w.db_version <= r.db_version)
    # writes did land while requests were being answered (not only before or after them)
    assert form == "per_batch" or mid_scan > 0


@pytest.mark.parametrize("form", ["serving", "per_batch"])
def test_requests_in_flight_compared_with_each_other(served, orc, form):
    """Uniqueness requests scanning together are compared with each other, as
    the reference compares a batch with itself (compare_query_against_self,
    mergeBatchResults): a request reports the earlier-admitted requests in
    flight with it (its batch, in the per-batch form) whose iris it matches.
    A burst of fresh irises (none in the DB) with noisy copies among them is
    admitted together; expected: the oracle's plaintext predicate between the
    request irises (31 rotations, threshold 0.375). A copy arriving after the
    burst has finished reports none."""
    eng, serving, synth = served
    R = serving.Request
    # (iris of dataset 78, code bits flipped): rids 0-5 one burst, 6 later
    plan = [(0, 0), (1, 0), (0, 1500), (1, 2500), (0, 3000), (2, 0), (0, 1000)]
    reqs, plain = [], []
    for k, (t, f) in enumerate(plan):
        sh, c, m = _query(synth, 78, t, 80 + k, f, 200 + k)
        plain.append((c, m))
        reqs.append(R("uniqueness", 0.0 if k < 6 else 0.5, shares=sh, rid=k))
    srv = serving.Server(eng, reserve_bytes=0) if form == "serving" else serving.PerBatchServer(eng, max_batch=8)
    out = {r.rid: r for r in srv.run(reqs)}
    for k in range(6):
        c, m = plain[k]
        want = [j for j in range(k) if orc.plain_match(c, m, plain[j][0][None], plain[j][1][None], 16384)[0]]
        assert out[k].batch_matches == want, f"{form}: request {k}"
        assert out[k].result == [], f"{form}: request {k} matched the DB"
    assert out[6].batch_matches == [] and out[6].result == []
    assert sum(len(out[k].batch_matches) for k in range(6)) >= 3, "the copies did not match their originals"
