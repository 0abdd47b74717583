"""DB-chunk records straight into HBM (InMemoryStore::load_single_record_from_s3,
iris-mpc-common/src/helpers/inmemory_store.rs:55-105; placement of
iris-mpc-store/src/loader.rs:196-240): a shuffled chunk with a gap and an id
above the cap loads the same DB as the u16 path, limbs and row sums (checked
through read-back and the share GEMM, share for share); bad chunks fail before
anything is written; a ServerActor fed from chunks finds stored irises."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


def _rand(rng, shape):
    return rng.integers(0, 1 << 16, shape, dtype=np.uint16)


def test_stored_records_equal_u16_load(irismpc):
    from irismpc import store

    rng = np.random.default_rng(9)
    N, Qr = 300, 62
    codes = irismpc.ShareDB(0, 512, Qr, 12800, ([1] * 8, [2] * 8))
    masks = irismpc.ShareDB(0, 512, Qr, 6400, ([3] * 8, [4] * 8))
    lc, rc = _rand(rng, (N, 12800)), _rand(rng, (N, 12800))
    lm, rm = _rand(rng, (N, 6400)), _rand(rng, (N, 6400))
    ids = rng.permutation(np.arange(1, N + 1))
    keep = (ids != 150) & (ids <= 290)  # a gap at row 149; ids above 290 are capped away
    chunk_ids = np.concatenate([ids[ids != 150]])
    sel = chunk_ids - 1
    chunk = store.encode_stored_records(chunk_ids, lc[sel], lm[sel], rc[sel], rm[sel])
    dbs = [codes.alloc_db(N), masks.alloc_db(N), codes.alloc_db(N), masks.alloc_db(N)]
    n, end = store.load_stored_records(codes, masks, *dbs, chunk, max_serial_id=290)
    assert n == int(keep.sum()) == 289 and end == 290

    want = [lc, lm, rc, rm]
    for d, w in zip(dbs, want):
        got = d.read_records(0, 290)
        rows = np.array([r for r in range(290) if r != 149])
        assert np.array_equal(got[rows], w[rows])

    # row sums: the GEMM over the stored DB equals the GEMM over a u16-loaded DB
    ref = codes.alloc_db(N)
    full = lc.copy()
    full[149] = 0x8080  # unloaded row: zero limbs, as in a fresh slice
    full[290:] = 0x8080
    ref.load_records(0, full)
    q = codes.upload_query_shares(_rand(rng, (2, 12800)))
    codes.dot_reduce(q, dbs[0], 320, 0, 1)
    a = codes.fetch_results(Qr * 320)
    codes.counter = 0
    codes.dot_reduce(q, ref, 320, 0, 1)
    b = codes.fetch_results(Qr * 320)
    assert np.array_equal(a, b)


def test_stored_records_rejected_before_writing(irismpc):
    from irismpc import store

    rng = np.random.default_rng(10)
    codes = irismpc.ShareDB(0, 256, 31, 12800, ([1] * 8, [2] * 8))
    masks = irismpc.ShareDB(0, 256, 31, 6400, ([3] * 8, [4] * 8))
    dbs = [codes.alloc_db(64), masks.alloc_db(64), codes.alloc_db(64), masks.alloc_db(64)]
    before = dbs[0].read_records(0, 4)

    def chunk(ids):
        n = len(ids)
        return store.encode_stored_records(ids, _rand(rng, (n, 12800)), _rand(rng, (n, 6400)),
                                           _rand(rng, (n, 12800)), _rand(rng, (n, 6400)))

    for ids, msg in (([1, 0, 3], "serial id 0"), ([1, 2, 1], "repeated"), ([2, 400], "beyond")):
        with pytest.raises(irismpc.IrisMpcError, match=msg):
            store.load_stored_records(codes, masks, *dbs, chunk(ids))
    assert np.array_equal(dbs[0].read_records(0, 4), before)
    assert store.load_stored_records(codes, masks, *dbs, b"") == (0, 0)


def test_server_from_stored_chunks(irismpc):
    from irismpc import store, synth
    from irismpc.server import BatchQuery, ServerActor

    N = 200
    plain = {e: [synth.synth_iris(ds, i) for i in range(N)] for e, ds in (("left", 41), ("right", 42))}
    sh = {}
    for k, e in enumerate(("left", "right")):
        sh[e] = synth.shares_of(np.array([c for c, _ in plain[e]]), np.array([m for _, m in plain[e]]), 80 + k)
    ids = np.arange(1, N + 1)[::-1]  # stored in reverse order
    chunks = [store.encode_stored_records(ids, sh["left"][p][0][ids - 1], sh["left"][p][1][ids - 1],
                                          sh["right"][p][0][ids - 1], sh["right"][p][1][ids - 1])
              for p in range(3)]
    actor = ServerActor(1024, 8)
    assert actor.load_stored_records(chunks) == [N, N, N]
    assert actor.current_db_size == N

    targets = [5, 77, 199]
    q = {e: [] for e in ("left", "right")}
    for j, t in enumerate(targets):
        for e in ("left", "right"):
            c, m = plain[e][t]
            q[e].append((synth.synth_flip(c, 60, j + (e == "right") * 10, 1000), m))
    qs = {}
    for k, e in enumerate(("left", "right")):
        qs[e] = synth.shares_of(np.array([c for c, _ in q[e]]), np.array([m for _, m in q[e]]), 90 + k)
    batch = BatchQuery([f"r{j}" for j in range(3)], ["uniqueness"] * 3, qs["left"], qs["right"])
    res = actor.process_batch_query(batch)
    assert res.merged_results == targets
