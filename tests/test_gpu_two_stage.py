"""Two-stage matching flow (iris-mpc-gpu/src/server/actor.rs:860-1131) on one
GPU: full scan of the left eye at 0.375, candidate subset (get_partial_results
+ OR-rule / reauth ids), both eyes compared on the subset at 0.345
(prefetch_subset + openResultsWithIndexMapping), AND / OR-rule merge.

test_two_stage_vs_plaintext: every output against a plaintext brute force with
the integer MPC predicate, per rotation: the candidate list, both eyes'
opened bitmaps, the (query, DB id, rotation) captures, the merged results,
counters and match lists; and the merged answer equals that of a full scan of
both eyes at 0.345. The queries cover: a match in both eyes; a left-eye
distance between 0.345 and 0.375 (a candidate that fails the second stage);
left-only and right-only matches under the OR rule and without it; a query
that is a rotated copy of its DB iris (rotation capture); a reauth target.

test_two_stage_multi_chunk_subset: 40,960 candidates (more than one subset
chunk of 32,768) on a 65,536-row DB of 64 distinct irises tiled; candidates,
final results and counters against the oracle's plaintext brute force.

test_anon_distance_capture: the anonymized-statistics distance capture of both
stages against the plaintext dot products.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_0375 = 16384


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


def _rotation_bits(orc, db_c, db_m, q_c, q_m, A):
    """[B][31][N] plaintext MPC predicate per query rotation."""
    N, B = len(db_c), len(q_c)
    bits = np.zeros((B, 31, N), bool)
    for j in range(B):
        rc, rm = orc.iris_all_rotations(q_c[j]), orc.iris_all_rotations(q_m[j])
        for t in range(31):
            mm = rm[t][None, :] & db_m
            M = np.bitwise_count(mm).sum(axis=1).astype(np.int64)
            H = np.bitwise_count((rc[t][None, :] ^ db_c) & mm).sum(axis=1).astype(np.int64)
            c = (M - 2 * H) & 0xFFFF
            r = (M * A - (c << 16) - 1) & 0xFFFFFFFF
            bits[j, t] = (r >> 31).astype(bool)
    return bits


def _tiled_matches(orc, ds, U, N, q_codes, q_masks, A):
    """[B][N] plaintext brute force of each query against a DB whose row r
    holds synthetic iris (ds, r mod U): the oracle's integer MPC predicate
    over all rotations against the U distinct irises, tiled."""
    db_c, db_m = map(np.array, zip(*[orc.synth_iris(ds, u) for u in range(U)]))
    reps = -(-N // U)
    return np.stack([np.tile(orc.plain_match(q_codes[j], q_masks[j], db_c, db_m, A), reps)[:N]
                     for j in range(len(q_codes))])


def _engines(irismpc, N, Qr, tag_base=0):
    base = [np.full(8, p + tag_base, np.uint32) for p in range(3)]
    ex = irismpc.Exchange(timeout_s=120)
    chunk = irismpc.DB_CHUNK_SIZE
    out = []
    for p in range(3):
        der = lambda n: (irismpc.derive_seed(base[p], n), irismpc.derive_seed(base[(p + 2) % 3], n))  # noqa: E731
        codes = irismpc.ShareDB(p, chunk, Qr, 12800, der(0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, Qr, 6400, der(1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, der(5), ex, tag=2)
        cmp = irismpc.DistanceComparator(Qr, N)
        out.append(irismpc.PartyEngines(codes, masks, circ, cmp, codes.alloc_db(N), masks.alloc_db(N)))
    return ex, out


def test_two_stage_vs_plaintext(irismpc, orc):
    from irismpc import synth
    from irismpc.two_stage import TwoStageMatcher

    N, B = 2000, 8
    Qr = B * 31
    rng = np.random.default_rng(77)
    t = rng.choice(N, size=8, replace=False)
    db = {}
    for e, ds in (("left", 3), ("right", 4)):
        c, m = zip(*[orc.synth_iris(ds, i) for i in range(N)])
        db[e] = (np.array(c), np.array(m))
    flips = {  # query j: (left target / flips, right target / flips); None = random iris
        0: ((t[0], 1280), (t[0], 1280)),
        1: ((t[1], 4608), (t[1], 1280)),   # left FHD ~0.36: candidate, fails 0.345
        2: ((t[2], 1280), None),           # left only, OR rule
        3: (None, (t[3], 1280)),           # right only, OR rule (candidate via the rule)
        4: ((t[4], 1280), None),           # left only, AND
        5: ("rot", "rot"),                 # rotated copy of DB iris t[5]
        6: (None, None),
        7: ((t[7], 2000), (t[7], 1500)),
    }
    q = {"left": ([], []), "right": ([], [])}
    for j in range(B):
        for side, e in enumerate(("left", "right")):
            spec = flips[j][side]
            dc, dm = db[e]
            if spec is None:
                c, m = orc.synth_iris(20 + side, j)
            elif spec == "rot":
                c = orc.synth_flip(orc.iris_all_rotations(dc[t[5]])[12], 40 + side, j, 640)
                m = orc.iris_all_rotations(dm[t[5]])[12]
            else:
                c, m = orc.synth_flip(dc[spec[0]], 30 + side, j, spec[1]), dm[spec[0]]
            q[e][0].append(c)
            q[e][1].append(m)
    q = {e: (np.array(v[0]), np.array(v[1])) for e, v in q.items()}
    or_rule = [[] for _ in range(B)]
    or_rule[2] = [int(t[2])]
    or_rule[3] = [int(t[3])]
    reauth = [int(t[6]), N + 5]  # the second is beyond the DB and skipped

    eyes, queries, keep = {}, {}, []
    for k, e in enumerate(("left", "right")):
        ex, parties = _engines(irismpc, N, Qr, tag_base=3 * k)
        keep.append(ex)
        db_sh = synth.shares_of(db[e][0], db[e][1], 50 + k)
        q_sh = synth.shares_of(q[e][0], q[e][1], 60 + k)
        for p in range(3):
            parties[p].code_db.load_full_db(db_sh[p][0])
            parties[p].mask_db.load_full_db(db_sh[p][1])
        eyes[e] = parties
        queries[e] = ([parties[p].codes.upload_query_shares(q_sh[p][0]) for p in range(3)],
                      [parties[p].masks.upload_query_shares(q_sh[p][1]) for p in range(3)])

    m = TwoStageMatcher(eyes, N, B, full_scan_side="left")
    res = m.match(queries, N, B, or_rule_indices=or_rule, reauth_targets=reauth)

    # plaintext answer
    l375 = _rotation_bits(orc, *db["left"], *q["left"], A_0375).any(axis=1)
    lr = _rotation_bits(orc, *db["left"], *q["left"], irismpc.THRESHOLD_A)
    rr = _rotation_bits(orc, *db["right"], *q["right"], irismpc.THRESHOLD_A)
    extra = {int(d) for ids in or_rule for d in ids} | {d for d in reauth if d < N}
    cand = np.array(sorted(set(np.nonzero(l375.any(axis=0))[0].tolist()) | extra), np.uint32)
    assert l375[1, t[1]] and not lr[1, :, t[1]].any(), "query 1 must fall between the thresholds"
    assert np.array_equal(res.candidates, cand)
    cmask = np.zeros(N, bool)
    cmask[cand] = True
    L, R = lr.any(axis=1) & cmask, rr.any(axis=1) & cmask
    or_pair = np.zeros((B, N), bool)
    for j, ids in enumerate(or_rule):
        or_pair[j, ids] = True
    merged = np.where(or_pair, L | R, L & R)
    full = np.where(or_pair, lr.any(axis=1) | rr.any(axis=1), lr.any(axis=1) & rr.any(axis=1))
    assert np.array_equal(merged, full), "two-stage answer differs from the full scan at 0.345"
    want = np.array([np.nonzero(r)[0].min() if r.any() else 0xFFFFFFFF for r in merged], np.uint32)
    assert want[0] == t[0] and want[2] == t[2] and want[3] == t[3] and want[5] == t[5] and want[7] == t[7]
    assert want[1] == want[4] == want[6] == 0xFFFFFFFF

    for p in range(3):
        assert np.array_equal(res.final_results[p], want), f"party {p} final results"
        assert np.array_equal(res.match_counters[p], merged.sum(axis=1)), f"party {p} match counters"
        for j in range(B):
            n = int(merged[j].sum())
            assert set(res.all_matches[p][j][:n].tolist()) == set(np.nonzero(merged[j])[0].tolist())
        for e, bits in (("left", L), ("right", R)):
            cnt, lists = res.partial[e][p]
            assert np.array_equal(cnt, bits.sum(axis=1)), f"party {p} {e} partial counters"
            for j in range(B):
                assert set(lists[j][:int(cnt[j])].tolist()) == set(np.nonzero(bits[j])[0].tolist())
        for e, rb in (("left", lr), ("right", rr)):
            exp = {(j, int(d), tt - 15) for j, tt, d in zip(*np.nonzero(rb & cmask[None, None, :]))}
            got = {(qi, di, r) for qi, dd in res.rotations[e][p].items() for di, rs in dd.items() for r in rs}
            n_got = sum(len(rs) for dd in res.rotations[e][p].values() for rs in dd.values())
            assert got == exp and n_got == len(exp), f"party {p} {e} rotation capture"
    assert (5, int(t[5]), 3) in {(qi, di, r) for qi, dd in res.rotations["left"][0].items()
                                 for di, rs in dd.items() for r in rs}


def test_two_stage_multi_chunk_subset(irismpc, orc):
    import torch

    from irismpc import synth
    from irismpc.two_stage import TwoStageMatcher

    N, U, B = 65536, 64, 64
    Qr = B * 31
    rng = np.random.default_rng(5)
    targets = rng.choice(U, size=40, replace=False)
    eyes, queries, keep, plain = {}, {}, [], {}
    for k, e in enumerate(("left", "right")):
        ex, parties = _engines(irismpc, N, Qr, tag_base=3 * k)
        keep.append(ex)
        ds = 70 + k
        synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, ds, ds + 30)
        # left eye: all 40 planted; right eye: the even ones
        planted = [(j, int(targets[j]), 1000 + 30 * j) for j in range(40) if e == "left" or j % 2 == 0]
        q_codes, q_masks, sh = synth.make_queries(B, planted, ds, ds + 20, ds + 10, ds + 40)
        plain[e] = (ds, q_codes, q_masks)
        eyes[e] = parties
        queries[e] = ([parties[p].codes.upload_query_shares(sh[p][0]) for p in range(3)],
                      [parties[p].masks.upload_query_shares(sh[p][1]) for p in range(3)])
    torch.cuda.synchronize()
    print("DBs built, matching", flush=True)
    res = TwoStageMatcher(eyes, N, B).match(queries, N, B)

    rows = np.arange(N, dtype=np.uint32)
    cand = rows[np.isin(rows % U, targets)]
    assert cand.size == 40 * N // U > irismpc.DB_CHUNK_SIZE
    assert np.array_equal(res.candidates, cand)
    # the oracle's brute force: candidates = rows some query matches in the
    # left eye at 0.375; final = first row matching in both eyes at 0.345
    pre = _tiled_matches(orc, plain["left"][0], U, N, *plain["left"][1:], A_0375)
    assert np.array_equal(rows[pre.any(axis=0)], cand)
    both = (_tiled_matches(orc, plain["left"][0], U, N, *plain["left"][1:], 20316) &
            _tiled_matches(orc, plain["right"][0], U, N, *plain["right"][1:], 20316))
    want = np.array([np.flatnonzero(both[j])[0] if both[j].any() else 0xFFFFFFFF for j in range(B)], np.uint32)
    cnt = both.sum(axis=1).astype(np.uint32)
    assert np.array_equal(want[0:40:2], targets[0:40:2]) and (cnt[0:40:2] == N // U).all()
    for p in range(3):
        assert np.array_equal(res.final_results[p], want), f"party {p} final results"
        assert np.array_equal(res.match_counters[p], cnt), f"party {p} match counters"
        lc, _ = res.partial["left"][p]
        assert np.array_equal(lc[:40], np.full(40, N // U)) and not lc[40:].any()
        left = res.rotations["left"][p]
        assert sorted(left) == list(range(40))
        for j in range(40):
            assert sorted(left[j]) == cand[(cand % U) == targets[j]].tolist()
            assert all(r == [0] for r in left[j].values())


def _rotation_dots(orc, db_c, db_m, q_c, q_m):
    """[B][31][N] plaintext mask popcount M and masked Hamming distance H."""
    N, B = len(db_c), len(q_c)
    M = np.zeros((B, 31, N), np.int64)
    H = np.zeros((B, 31, N), np.int64)
    for j in range(B):
        rc, rm = orc.iris_all_rotations(q_c[j]), orc.iris_all_rotations(q_m[j])
        for t in range(31):
            mm = rm[t][None, :] & db_m
            M[j, t] = np.bitwise_count(mm).sum(axis=1)
            H[j, t] = np.bitwise_count((rc[t][None, :] ^ db_c) & mm).sum(axis=1)
    return M, H


def test_anon_distance_capture(irismpc, orc):
    """Match-distance capture of the anonymized statistics (openResults /
    storeAnonStatsWithIndexMapping, dot/kernel.cu:54-210): the full-scan eye
    captures every (DB row, query row) opened at 0.375, the other eye those in
    the candidate subset (its separate 0.375 pass), both plus every rotation of
    the reauth target pairs; the captured shares of the three parties add up
    to the plaintext code dot M - 2H and mask dot M, with replicated b shares,
    and match ids encode (batch, DB row, query row)."""
    from irismpc import synth
    from irismpc.anon_stats import DistanceBuffers
    from irismpc.two_stage import TwoStageMatcher

    N, B = 1500, 8
    Qr = B * 31
    rng = np.random.default_rng(78)
    t = rng.choice(N, size=4, replace=False)
    db = {}
    for e, ds in (("left", 5), ("right", 6)):
        c, m = zip(*[orc.synth_iris(ds, i) for i in range(N)])
        db[e] = (np.array(c), np.array(m))
    q = {"left": ([], []), "right": ([], [])}
    plan = {0: (1280, 1280), 1: (4608, 1280), 2: (1280, None), 3: (None, 1280)}
    for j in range(B):
        for side, e in enumerate(("left", "right")):
            f = plan.get(j, (None, None))[side]
            if f is None:
                c, m = orc.synth_iris(25 + side, j)
            else:
                c, m = orc.synth_flip(db[e][0][t[j]], 35 + side, j, f), db[e][1][t[j]]
            q[e][0].append(c)
            q[e][1].append(m)
    q = {e: (np.array(v[0]), np.array(v[1])) for e, v in q.items()}
    reauth_db, reauth_q = int(t[0]), 5  # query 5 (random) asks for DB row t[0]: captured, no match

    eyes, queries, keep = {}, {}, []
    for k, e in enumerate(("left", "right")):
        ex, parties = _engines(irismpc, N, Qr, tag_base=3 * k)
        keep.append(ex)
        db_sh = synth.shares_of(db[e][0], db[e][1], 51 + k)
        q_sh = synth.shares_of(q[e][0], q[e][1], 61 + k)
        for p in range(3):
            parties[p].code_db.load_full_db(db_sh[p][0])
            parties[p].mask_db.load_full_db(db_sh[p][1])
        eyes[e] = parties
        queries[e] = ([parties[p].codes.upload_query_shares(q_sh[p][0]) for p in range(3)],
                      [parties[p].masks.upload_query_shares(q_sh[p][1]) for p in range(3)])
    bufs = {e: [DistanceBuffers(1 << 14) for _ in range(3)] for e in ("left", "right")}
    batch_id = 7
    anon = {"buffers": bufs, "batch_id": batch_id, "targets": [reauth_db * B + reauth_q]}
    res = TwoStageMatcher(eyes, N, B).match(queries, N, B, reauth_targets=[reauth_db], anon=anon)

    cand = set(res.candidates.tolist())
    assert reauth_db in cand
    for e in ("left", "right"):
        M, H = _rotation_dots(orc, *db[e], *q[e])
        c = (M - 2 * H) & 0xFFFF
        r = (M * 16384 - (c << 16) - 1) & 0xFFFFFFFF
        match375 = (r >> 31).astype(bool)  # [B][31][N]
        want = set()
        for j, tt, d in zip(*np.nonzero(match375)):
            if e == "left" or d in cand:
                want.add((int(d), int(j * 31 + tt)))
        want |= {(reauth_db, reauth_q * 31 + tt) for tt in range(31)}
        got, shares = None, []
        for p in range(3):
            n, idx, ca, cb, ma, mb = bufs[e][p].fetch()
            assert n == len(idx) and n < bufs[e][p].capacity
            order = np.argsort(idx)
            shares.append((idx[order], ca[order], cb[order], ma[order], mb[order]))
            pairs = {(int(i % (N * Qr)) // Qr, int(i % Qr)) for i in idx}
            assert all(int(i) // (N * Qr) == batch_id for i in idx)
            assert got is None or pairs == got
            got = pairs
        assert got == want, f"{e} captured pairs"
        assert len(shares[0][0]) == len(want)
        for p in range(3):
            assert np.array_equal(shares[p][0], shares[0][0])
            assert np.array_equal(shares[p][2], shares[(p + 2) % 3][1])  # b = previous party's a
            assert np.array_equal(shares[p][4], shares[(p + 2) % 3][3])
        code = (shares[0][1].astype(np.int64) + shares[1][1] + shares[2][1]) & 0xFFFF
        mask = (shares[0][3].astype(np.int64) + shares[1][3] + shares[2][3]) & 0xFFFF
        for k, i in enumerate(shares[0][0]):
            d, row = int(i % (N * Qr)) // Qr, int(i % Qr)
            j, tt = divmod(row, 31)
            assert code[k] == (M[j, tt, d] - 2 * H[j, tt, d]) & 0xFFFF
            assert mask[k] == M[j, tt, d] & 0xFFFF


def test_pipelined_scan_partial_last_chunk(irismpc, orc):
    """The chunk loop with two result slots (GEMMs of chunk c + 1 beside the
    circuits of chunk c) over four chunks, the last one partial, equals a
    chunk-by-chunk scan and the oracle's plaintext brute force, bit for bit."""
    import torch

    from irismpc import synth

    N, U, B = 3 * 32768 + 1000, 64, 16
    Qr = B * 31
    rng = np.random.default_rng(9)
    targets = rng.choice(U, size=10, replace=False)
    ex, parties = _engines(irismpc, N, Qr, tag_base=6)
    synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, 80, 110)
    planted = [(j, int(targets[j]), 1000 + 50 * j) for j in range(10)]
    q_codes, q_masks, sh = synth.make_queries(B, planted, 80, 100, 90, 120)
    qc = [parties[p].codes.upload_query_shares(sh[p][0]) for p in range(3)]
    qm = [parties[p].masks.upload_query_shares(sh[p][1]) for p in range(3)]
    streams = [irismpc.stream_create() for _ in range(3)]
    rows = np.arange(N)
    want = _tiled_matches(orc, 80, U, N, q_codes, q_masks, A_0375)
    for j, t, _ in planted:  # the planted copies are among the matches
        assert want[j][rows % U == t].all()

    def bits(bm):
        return np.unpackbits(bm.cpu().numpy().view(np.uint8), bitorder="little")[: B * N].reshape(B, N)

    maps = []
    for pipelined in (True, False):
        bms = [torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
        torch.cuda.synchronize()
        ptrs = [b.data_ptr() for b in bms]
        if pipelined:
            irismpc.scan_db_trio(parties, qc, qm, N, B, irismpc.THRESHOLD_ANON_STATS_A, ptrs, streams)
        else:
            for off in range(0, N, irismpc.DB_CHUNK_SIZE):
                irismpc.scan_chunk_trio(parties, qc, qm, off, N, B, irismpc.THRESHOLD_ANON_STATS_A, ptrs, streams)
        for s in streams:
            irismpc.stream_synchronize(s)
        got = [bits(b) for b in bms]
        assert all(np.array_equal(g, got[0]) for g in got), "parties opened different bits"
        maps.append(got[0])
    assert np.array_equal(maps[0], maps[1])
    assert np.array_equal(maps[0].astype(bool), want)


def test_multi_chunk_step_equals_single_steps(irismpc, orc):
    """Several chunks in one step (dot shares side by side, one reshare and one
    comparison) open the same bits as one step per chunk: four chunks with a
    partial last one, a query count whose rows need 512-row padding, and a
    step that runs past the DB's end."""
    import torch

    from irismpc import synth

    N, U, B = 3 * 32768 + 1000, 64, 5
    rows = B * 31
    rng = np.random.default_rng(10)
    targets = rng.choice(U, size=3, replace=False)
    ex, parties = _engines(irismpc, N, 4 * rows, tag_base=9)  # result buffers for four chunks of `rows`
    synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, 90, 120)
    planted = [(j, int(targets[j]), 900 + 100 * j) for j in range(3)]
    q_codes, q_masks, sh = synth.make_queries(B, planted, 90, 110, 100, 130)
    qc = [parties[p].codes.upload_query_shares_rows(sh[p][0], rows) for p in range(3)]
    qm = [parties[p].masks.upload_query_shares_rows(sh[p][1], rows) for p in range(3)]
    streams = [irismpc.stream_create() for _ in range(3)]
    A = irismpc.THRESHOLD_ANON_STATS_A
    want = _tiled_matches(orc, 90, U, N, q_codes, q_masks, A)
    for j, t, _ in planted:  # the planted copies are among the matches
        assert want[j][np.arange(N) % U == t].all()

    def run(steps):
        bms = [torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
        torch.cuda.synchronize()
        ptrs = [b.data_ptr() for b in bms]
        for off, k in steps:
            irismpc.scan_chunks_trio(parties, qc, qm, off, k, N, B, A, ptrs, streams)
        for s in streams:
            irismpc.stream_synchronize(s)
        got = [np.unpackbits(b.cpu().numpy().view(np.uint8), bitorder="little")[: B * N].reshape(B, N)
               for b in bms]
        assert all(np.array_equal(g, got[0]) for g in got), "parties opened different bits"
        return got[0].astype(bool)

    C = irismpc.DB_CHUNK_SIZE
    single = run([(c * C, 1) for c in range(4)])
    assert np.array_equal(single, want)
    assert np.array_equal(run([(0, 4)]), single)
    assert np.array_equal(run([(0, 2), (2 * C, 4)]), single)  # the second step stops at the DB's end
    assert np.array_equal(run([(0, 8)]), single)  # a lone query's step size, past the DB's end


def _flip_first(code, n):
    """IrisCodeArray::flip_bit(i) for i in 0..n (bit i = bit i % 64 of word i / 64)."""
    out = np.array(code, np.uint64).copy()
    full, rest = divmod(n, 64)
    out[:full] ^= np.uint64(0xFFFFFFFFFFFFFFFF)
    if rest:
        out[full] ^= np.uint64((1 << rest) - 1)
    return out


def test_close_to_threshold_two_stage(irismpc, orc):
    """The reference's CloseToThreshold / flipped-code cases
    (iris-mpc-common/src/test.rs:786-815, 1181-1220): a full-mask DB iris with
    its first THRESHOLD_ABSOLUTE + v code bits flipped (THRESHOLD_ABSOLUTE =
    4,416 = 0.345 x 12,800, v in -99..99) in each eye, through the share GEMM,
    the circuits and the two-stage flow; plus left eyes at the 0.375 boundary
    (4,800 + v flips), which decide the candidate list. With full masks the
    integer predicate flips exactly between v = 0 and v = 1; the plaintext
    brute force must agree with that closed form, and every party's merged
    results, counters and candidates with both."""
    from irismpc import synth
    from irismpc.two_stage import TwoStageMatcher

    N, B = 1024, 32
    Qr = B * 31
    T345, T375 = 12800 * 345 // 1000, 12800 * 375 // 1000
    rng = np.random.default_rng(1181)
    ones = np.full(200, np.uint64(0xFFFFFFFFFFFFFFFF))
    db = {e: (np.array([orc.synth_iris(ds, i)[0] for i in range(N)]), np.tile(ones, (N, 1)))
          for e, ds in (("left", 5), ("right", 6))}
    edge = [(0, 0), (-1, -1), (1, 0), (0, 1), (1, 1), (-1, 1), (-99, -99), (99, -99), (-99, 99),
            (-50, -3), (3, -50), (0, -1)]
    spec = [(T345 + a, T345 + b) for a, b in edge]
    spec += [(T375 + v, T345 - 20) for v in (0, -1, 1, 2)]  # 0.375 boundary in the full-scan eye
    spec += [(T345 + int(a), T345 + int(b)) for a, b in rng.integers(-99, 100, (B - len(spec), 2))]
    t = rng.choice(N, size=B, replace=False)
    q = {e: (np.array([_flip_first(db[e][0][t[j]], spec[j][k]) for j in range(B)]), np.tile(ones, (B, 1)))
         for k, e in enumerate(("left", "right"))}

    eyes, queries, keep = {}, {}, []
    for k, e in enumerate(("left", "right")):
        ex, parties = _engines(irismpc, N, Qr, tag_base=20 + 3 * k)
        keep.append(ex)
        db_sh = synth.shares_of(db[e][0], db[e][1], 70 + k)
        q_sh = synth.shares_of(q[e][0], q[e][1], 80 + k)
        for p in range(3):
            parties[p].code_db.load_full_db(db_sh[p][0])
            parties[p].mask_db.load_full_db(db_sh[p][1])
        eyes[e] = parties
        queries[e] = ([parties[p].codes.upload_query_shares(q_sh[p][0]) for p in range(3)],
                      [parties[p].masks.upload_query_shares(q_sh[p][1]) for p in range(3)])
    res = TwoStageMatcher(eyes, N, B, full_scan_side="left").match(queries, N, B)

    # plaintext brute force at both thresholds, and the closed form it must agree with
    plain = {(e, A): np.array([orc.plain_match(q[e][0][j], q[e][1][j], *db[e], A) for j in range(B)])
             for e in ("left", "right") for A in (A_0375, irismpc.THRESHOLD_A)}
    for j, (fl, fr) in enumerate(spec):
        assert plain["left", A_0375][j, t[j]] == (fl <= T375)
        assert plain["left", irismpc.THRESHOLD_A][j, t[j]] == (fl <= T345)
        assert plain["right", irismpc.THRESHOLD_A][j, t[j]] == (fr <= T345)
    cand = np.flatnonzero(plain["left", A_0375].any(axis=0)).astype(np.uint32)
    assert np.array_equal(res.candidates, cand)
    cmask = np.zeros(N, bool)
    cmask[cand] = True
    merged = plain["left", irismpc.THRESHOLD_A] & plain["right", irismpc.THRESHOLD_A] & cmask[None, :]
    want = np.array([np.flatnonzero(r).min() if r.any() else 0xFFFFFFFF for r in merged], np.uint32)
    assert sum(w != 0xFFFFFFFF for w in want) >= 4 and sum(w == 0xFFFFFFFF for w in want) >= 8
    for p in range(3):
        assert np.array_equal(res.final_results[p], want), f"party {p} final results"
        assert np.array_equal(res.match_counters[p], merged.sum(axis=1)), f"party {p} match counters"
