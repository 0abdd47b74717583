"""End-to-end three-party DB scan on one GPU vs the oracle and plaintext.

Config-1 shape (scaled for test time): synthetic iris codes with paired masks,
planted queries (DB copies with flipped bits), GR shares for three parties,
query preprocessing on the GPU, then the chunk loop of
compare_query_against_db_and_self_prefiltering (iris-mpc-gpu/src/server/
actor.rs:2246-2589): share GEMM + zero-share masks, reshare, threshold circuit
at 0.375 (A = 16384), open, OR over rotations into the match bitmap.
Checked: (1) every party's bitmap equals the plaintext brute force with the
integer MPC predicate, (2) the masked dot-product shares and (3) the circuit
result shares equal the oracle's, share for share.
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


def _workload(orc, N, B, planted, flips, seed_ds=1):
    db_c, db_m = zip(*[orc.synth_iris(seed_ds, i) for i in range(N)])
    db_c, db_m = np.array(db_c), np.array(db_m)
    rng = np.random.default_rng(N + B)
    targets = rng.choice(N, size=len(planted), replace=False)
    q_c, q_m = [], []
    for j in range(B):
        if j in planted:
            t = int(targets[planted.index(j)])
            q_c.append(orc.synth_flip(db_c[t], seed_ds + 10, j, flips))
            q_m.append(db_m[t])
        else:
            c, m = orc.synth_iris(seed_ds + 20, j)
            q_c.append(c)
            q_m.append(m)
    return db_c, db_m, np.array(q_c), np.array(q_m)


def _shares(orc, codes, masks, ds):
    cs = np.zeros((3, len(codes), 12800), np.uint16)
    ms = np.zeros((3, len(codes), 6400), np.uint16)
    for i, (c, m) in enumerate(zip(codes, masks)):
        a, b = orc.synth_shares(c, m, ds, i)
        cs[:, i], ms[:, i] = a, b
    return cs, ms


def _plain_bitmap(orc, db_c, db_m, q_c, q_m, A):
    N, B = len(db_c), len(q_c)
    bits = np.zeros(B * N, bool)
    for j in range(B):
        rc, rm = orc.iris_all_rotations(q_c[j]), orc.iris_all_rotations(q_m[j])
        for t in range(31):
            mm = rm[t][None, :] & db_m
            M = np.bitwise_count(mm).sum(axis=1).astype(np.int64)
            H = np.bitwise_count((rc[t][None, :] ^ db_c) & mm).sum(axis=1).astype(np.int64)
            c = (M - 2 * H) & 0xFFFF
            r = (M * A - (c << 16) - 1) & 0xFFFFFFFF
            bits[j * N:(j + 1) * N] |= (r >> 31).astype(bool)
    return bits


@pytest.mark.parametrize("N,B,planted,flips", [
    (2000, 8, [0, 2, 5], 1280),
    # the CPU-runnable configuration as stated: DB 4,096, 32 queries x 31
    # rotations, 8 of them DB copies with exactly 1,920 (15 %) code bits flipped
    (4096, 32, [0, 4, 8, 12, 16, 20, 24, 28], 1920),
], ids=["small", "config1"])
def test_pipeline_three_party(irismpc, orc, N, B, planted, flips):
    import os

    import torch

    orc.set_threads(min(16, os.cpu_count() or 1))
    Qr = B * 31
    db_c, db_m, q_c, q_m = _workload(orc, N, B, planted, flips=flips)
    dbc_sh, dbm_sh = _shares(orc, db_c, db_m, 30)
    qc_sh, qm_sh = _shares(orc, q_c, q_m, 31)

    base = [np.full(8, p, np.uint32) for p in range(3)]
    seeds = [(base[p], base[(p + 2) % 3]) for p in range(3)]
    der = lambda p, n: (irismpc.derive_seed(seeds[p][0], n), irismpc.derive_seed(seeds[p][1], n))  # noqa: E731
    ex = irismpc.Exchange(timeout_s=120)
    chunk = irismpc.DB_CHUNK_SIZE
    parties, qcs, qms, bitmaps, streams = [], [], [], [], []
    for p in range(3):
        codes = irismpc.ShareDB(p, chunk, Qr, 12800, der(p, 0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, Qr, 6400, der(p, 1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, der(p, 5), ex, tag=2)
        cmp = irismpc.DistanceComparator(Qr, N)
        cdb, mdb = codes.alloc_db(N), masks.alloc_db(N)
        cdb.load_full_db(dbc_sh[p])
        mdb.load_full_db(dbm_sh[p])
        parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, cdb, mdb))
        qcs.append(codes.upload_query_shares(qc_sh[p]))
        qms.append(masks.upload_query_shares(qm_sh[p]))
        bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
        streams.append(irismpc.stream_create())
    torch.cuda.synchronize()
    irismpc.scan_db_trio(parties, qcs, qms, N, B, A_0375, [b.data_ptr() for b in bitmaps], streams)
    for s in streams:
        irismpc.stream_synchronize(s)

    exp = _plain_bitmap(orc, db_c, db_m, q_c, q_m, A_0375)
    assert exp.sum() >= len(planted)
    for j in planted:  # every planted copy is found
        assert exp[j * N:(j + 1) * N].any()
    for p in range(3):
        got = np.unpackbits(bitmaps[p].cpu().numpy().view(np.uint8), bitorder="little")[:N * B]
        assert np.array_equal(got.astype(bool), exp), f"party {p} bitmap"

    # share-level parity of the masked dot products (single chunk: results buffer)
    dot_chunk = (N + 63) // 64 * 64
    rows_c = [np.zeros((Qr, 12800), np.uint16) for _ in range(3)]
    rows_m = [np.zeros((Qr, 6400), np.uint16) for _ in range(3)]
    for p in range(3):
        for j in range(B):
            rows_c[p][j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qc_sh[p, j]))
            rows_m[p][j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qm_sh[p, j]))
    pad_c = np.zeros((dot_chunk, 12800), np.uint16)
    pad_m = np.zeros((dot_chunk, 6400), np.uint16)
    code_sh, mask_sh = [], []
    for p in range(3):
        pad_c[:] = 0x8080  # zero limbs of rows beyond the DB
        pad_m[:] = 0x8080
        pad_c[:N], pad_m[:N] = dbc_sh[p], dbm_sh[p]
        cd = orc.u16_dot(rows_c[p], pad_c).reshape(-1)
        md = orc.u16_dot(rows_m[p], pad_m).reshape(-1)
        sc, sm = der(p, 0), der(p, 1)
        code_sh.append(orc.mask_dot_share(cd, 1, sc[0], sc[1], 0, 0))
        mask_sh.append(orc.mask_dot_share(md, 2, sm[0], sm[1], 0, 0))
        got_c = parties[p].codes.fetch_results(Qr * dot_chunk)
        got_m = parties[p].masks.fetch_results(Qr * dot_chunk)
        assert np.array_equal(got_c, code_sh[p]), f"party {p} code dot shares"
        assert np.array_equal(got_m, mask_sh[p]), f"party {p} mask dot shares"
    # circuit result shares: replicated (a = own, b = prev's) of the dot shares
    ca = code_sh
    cb = [code_sh[(p + 2) % 3] for p in range(3)]
    ma = mask_sh
    mb = [mask_sh[(p + 2) % 3] for p in range(3)]
    so = np.array([der(p, 5)[0] for p in range(3)])
    sp = np.array([der(p, 5)[1] for p in range(3)])
    ra, rb, _ = orc.threshold(ca, cb, ma, mb, Qr * chunk // 64, A_0375, so, sp)
    for p in range(3):
        a, b, _c = parties[p].circuits.open_views()
        W = Qr * dot_chunk // 64
        assert np.array_equal(irismpc.download(a, W, np.uint64), ra[p]), f"party {p} result a"
        assert np.array_equal(irismpc.download(b, W, np.uint64), rb[p]), f"party {p} result b"
    orc.set_threads(1)


def test_pipeline_multi_chunk_share_parity(irismpc, orc):
    """Three chunks (two full, the last 1,000 rows) in the chunk loop with two
    result slots: every bitmap bit equals the plaintext brute force, and the
    last chunk's masked dot shares and circuit result shares equal the
    oracle's at the counters that the two earlier chunks advanced to (ShareDB:
    mask + OTP blocks per chunk; Circuits: one schedule per chunk)."""
    import os

    import torch

    from irismpc import synth

    orc.set_threads(min(16, os.cpu_count() or 1))
    chunk = irismpc.DB_CHUNK_SIZE
    N, U, B = 2 * chunk + 1000, 1024, 2
    Qr = B * 31
    targets = [517, 999]  # copies in every chunk, the last included
    dc, dm = zip(*[orc.synth_iris(41, u) for u in range(U)])
    dc, dm = np.array(dc), np.array(dm)
    q_c = np.array([orc.synth_flip(dc[t], 42, j, 1500) for j, t in enumerate(targets)])
    q_m = dm[targets]
    qc_sh, qm_sh = _shares(orc, q_c, q_m, 43)

    base = [np.full(8, 10 + p, np.uint32) for p in range(3)]
    der = lambda p, n: (irismpc.derive_seed(base[p], n), irismpc.derive_seed(base[(p + 2) % 3], n))  # noqa: E731
    ex = irismpc.Exchange(timeout_s=120)
    parties, qcs, qms, bitmaps, streams = [], [], [], [], []
    for p in range(3):
        codes = irismpc.ShareDB(p, chunk, Qr, 12800, der(p, 0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, Qr, 6400, der(p, 1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, der(p, 5), ex, tag=2)
        cmp = irismpc.DistanceComparator(Qr, N)
        parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, codes.alloc_db(N), masks.alloc_db(N)))
        bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
        streams.append(irismpc.stream_create())
    synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, 41, 44)
    for p in range(3):
        qcs.append(parties[p].codes.upload_query_shares(qc_sh[p]))
        qms.append(parties[p].masks.upload_query_shares(qm_sh[p]))
    torch.cuda.synchronize()
    irismpc.scan_db_trio(parties, qcs, qms, N, B, A_0375, [b.data_ptr() for b in bitmaps], streams)
    for s in streams:
        irismpc.stream_synchronize(s)

    hit = np.array([orc.plain_match(q_c[j], q_m[j], dc, dm, A_0375) for j in range(B)])  # [B][U]
    exp = np.concatenate([np.tile(hit[j], -(-N // U))[:N] for j in range(B)])
    assert all(hit[j, t] for j, t in enumerate(targets))
    for p in range(3):
        got = np.unpackbits(bitmaps[p].cpu().numpy().view(np.uint8), bitorder="little")[:N * B]
        assert np.array_equal(got.astype(bool), exp), f"party {p} bitmap"

    # counters after the two full chunks (rows % 8 != 0: chunks are padded to 512)
    last0, real = 2 * chunk, N - 2 * chunk
    dot_chunk = -(-real // 512) * 512
    blocks = lambda width: -(-(width * Qr * 2) // 64)  # noqa: E731
    share_ctr = 2 * 2 * blocks(chunk)
    assert parties[0].codes.counter == share_ctr + 2 * blocks(dot_chunk)
    zc = [np.zeros(Qr * chunk, np.uint16)] * 3
    so = np.array([der(p, 5)[0] for p in range(3)])
    sp = np.array([der(p, 5)[1] for p in range(3)])
    alloc = Qr * chunk // 64
    circ_ctr = 0
    for _ in range(2):  # the schedule's advance depends on the geometry only
        _, _, circ_ctr = orc.threshold(zc, zc, zc, zc, alloc, A_0375, so, sp, counter=circ_ctr)

    rows_c = [np.zeros((Qr, 12800), np.uint16) for _ in range(3)]
    rows_m = [np.zeros((Qr, 6400), np.uint16) for _ in range(3)]
    code_sh, mask_sh = [], []
    for p in range(3):
        for j in range(B):
            rows_c[p][j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qc_sh[p, j]))
            rows_m[p][j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qm_sh[p, j]))
        pad_c = np.full((dot_chunk, 12800), 0x8080, np.uint16)  # zero limbs beyond the DB
        pad_m = np.full((dot_chunk, 6400), 0x8080, np.uint16)
        pad_c[:real] = parties[p].code_db.read_records(last0, real)
        pad_m[:real] = parties[p].mask_db.read_records(last0, real)
        sc, sm = der(p, 0), der(p, 1)
        code_sh.append(orc.mask_dot_share(orc.u16_dot(rows_c[p], pad_c).reshape(-1), 1, sc[0], sc[1], 0, share_ctr))
        mask_sh.append(orc.mask_dot_share(orc.u16_dot(rows_m[p], pad_m).reshape(-1), 2, sm[0], sm[1], 0, share_ctr))
        assert np.array_equal(parties[p].codes.fetch_results(Qr * dot_chunk), code_sh[p]), f"party {p} code shares"
        assert np.array_equal(parties[p].masks.fetch_results(Qr * dot_chunk), mask_sh[p]), f"party {p} mask shares"
    ra, rb, end = orc.threshold(code_sh, [code_sh[(p + 2) % 3] for p in range(3)], mask_sh,
                                [mask_sh[(p + 2) % 3] for p in range(3)], alloc, A_0375, so, sp, counter=circ_ctr)
    W = Qr * dot_chunk // 64
    for p in range(3):
        a, b, _c = parties[p].circuits.open_views()
        assert np.array_equal(irismpc.download(a, W, np.uint64), ra[p]), f"party {p} result a"
        assert np.array_equal(irismpc.download(b, W, np.uint64), rb[p]), f"party {p} result b"
        assert parties[p].circuits.counter == end
    orc.set_threads(1)


def test_pipeline_config1_golden(irismpc, orc):
    """The CPU-runnable configuration as stated (4,096 DB irises, 32 queries x
    31 rotations, 8 copies with exactly 1,920 flips, targets from the ChaCha20
    key-0 stream) through the three-party scan: the bitmap, the first DB
    columns of every party's masked dot shares and the first result words of
    its circuit shares equal the committed golden vectors."""
    import os
    import sys

    import torch

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_config1 as g

    want = g.load()
    db_c, db_m, q_c, q_m, targets = g.inputs(orc)
    assert targets == want["targets"].tolist()
    dbc_sh, dbm_sh = g.shares(orc, db_c, db_m, 30)
    qc_sh, qm_sh = g.shares(orc, q_c, q_m, 31)
    N, B, Qr, chunk = g.N, g.B, g.B * 31, irismpc.DB_CHUNK_SIZE
    ex = irismpc.Exchange(timeout_s=120)
    parties, qcs, qms, bitmaps, streams = [], [], [], [], []
    for p in range(3):
        codes = irismpc.ShareDB(p, chunk, Qr, 12800, g.seeds(orc, p, 0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, Qr, 6400, g.seeds(orc, p, 1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, g.seeds(orc, p, 5), ex, tag=2)
        cmp = irismpc.DistanceComparator(Qr, N)
        cdb, mdb = codes.alloc_db(N), masks.alloc_db(N)
        cdb.load_full_db(dbc_sh[p])
        mdb.load_full_db(dbm_sh[p])
        parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, cdb, mdb))
        qcs.append(codes.upload_query_shares(qc_sh[p]))
        qms.append(masks.upload_query_shares(qm_sh[p]))
        bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
        streams.append(irismpc.stream_create())
    torch.cuda.synchronize()
    irismpc.scan_db_trio(parties, qcs, qms, N, B, g.A, [b.data_ptr() for b in bitmaps], streams)
    for s in streams:
        irismpc.stream_synchronize(s)
    for p in range(3):
        got = bitmaps[p].cpu().numpy().view(np.uint64)[:N * B // 64]
        assert np.array_equal(got, want["bitmap"]), f"party {p} bitmap"
        cd = parties[p].codes.fetch_results(Qr * N).reshape(Qr, N)[:, :g.COLS]
        md = parties[p].masks.fetch_results(Qr * N).reshape(Qr, N)[:, :g.COLS]
        assert np.array_equal(cd, want["code_dot_cols"][p]), f"party {p} code dot shares"
        assert np.array_equal(md, want["mask_dot_cols"][p]), f"party {p} mask dot shares"
        a, b, _c = parties[p].circuits.open_views()
        assert np.array_equal(irismpc.download(a, g.WORDS, np.uint64), want["result_a"][p]), f"party {p} result a"
        assert np.array_equal(irismpc.download(b, g.WORDS, np.uint64), want["result_b"][p]), f"party {p} result b"
