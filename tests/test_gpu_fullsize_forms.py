"""The batching and serving forms at their stated sizes, against the oracle.

test_c4_batching_fullsize: one party's code and mask DB shares of 4,194,304
rows (161 GB); a seeded subset of the 200 log-uniform batches of 1-256
queries (the first seven and the largest). The packed launches replayed from
one HIP graph give every batch exactly its plain per-batch shares, and the
first and last chunk of the largest and the smallest batch equal the
oracle's rotated-row u16 GEMM with its ChaCha12 zero-share masks at the
counters of the plain schedule.

test_c5_serving_fullsize: three parties' DB shares of 2,097,152 rows (one
eye, 32,768 distinct synthetic irises tiled) and a Poisson mix of
uniqueness, reauth, insertion and deletion requests through the request
loop. Every uniqueness and reauth result equals the plaintext brute force
with the integer MPC predicate over the DB as the request saw it (rows
deleted during the run excepted), and two requests give the same matches
alone as in the mix.
"""
import gc

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


def _free_device(torch):
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _rows(orc, party, shares):
    out = np.zeros((shares.shape[0] * 31, shares.shape[1]), np.uint16)
    for j in range(shares.shape[0]):
        out[j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(party, shares[j]))
    return out


def test_c4_batching_fullsize(irismpc, orc):
    import os

    import torch

    from irismpc.batching import BatchRunner, Matrix, counter_advance, log_uniform_batch_sizes, range_advance

    _free_device(torch)
    threads = min(16, os.cpu_count() or 1)
    orc.set_threads(threads)
    N, chunk = 4_194_304, irismpc.DB_CHUNK_SIZE
    all_sizes = log_uniform_batch_sizes(200, seed=4)
    pick = sorted(set(range(7)) | {int(np.argmax(all_sizes))})
    sizes = [int(all_sizes[i]) for i in pick]
    rng = np.random.default_rng(5)
    ex = irismpc.Exchange()
    seeds = (irismpc.derive_seed([0] * 8, 0), irismpc.derive_seed([2] * 8, 0))
    mats = []
    for tag, (K, mult) in enumerate(((12800, 1), (6400, 2))):
        eng = irismpc.ShareDB(0, chunk, 256 * 31, K, seeds, comm=ex, tag=tag)
        db = eng.alloc_db(N)
        db.fill_random(0, N, np.arange(8) + 100 * tag, 0)
        mats.append(Matrix(eng, db, N, mult, [rng.integers(0, 1 << 16, (b, K), dtype=np.uint16) for b in sizes]))
    stream = irismpc.stream_create()
    runner = BatchRunner(mats, sizes, chunk, stream, max_queries=256)
    runner.run_plain(ctr0=0)
    plain = [[o.clone() for o in outs] for outs in runner.outs]
    for outs in runner.outs:
        for o in outs:
            o.zero_()
    torch.cuda.synchronize()
    runner.run_packed(ctr0=0, graph=True)
    torch.cuda.synchronize()
    for k in range(2):
        for i, b in enumerate(sizes):
            n = chunk * b * 31
            assert torch.equal(runner.outs[k][i][:n], plain[k][i][:n]), f"matrix {k} batch {i} packed != plain"

    big, small = int(np.argmax(sizes)), int(np.argmin(sizes))
    cols = 2048
    last0 = N - chunk
    for k, m in enumerate(mats):
        base = np.cumsum([0] + [range_advance(N, chunk, b * 31) for b in sizes])
        for i in (big, small):
            rows = _rows(orc, 0, m.shares[i])
            qr = rows.shape[0]
            # chunk 0 (its own launch at the plain schedule's counter) and the last chunk (the plain run's)
            tmp = torch.zeros(chunk * qr + 64, dtype=torch.int16, device="cuda")
            m.engine.counter = int(base[i])
            q = m.engine.upload_query_shares_rows(m.shares[i], sizes[i] * 31)
            m.engine.dot_reduce_range(q, m.db, 0, chunk, chunk, m.mult, tmp.data_ptr(), stream)
            irismpc.stream_synchronize(stream)
            for c0, got_t, ctr in ((0, tmp, int(base[i])),
                                   (last0, plain[k][i], int(base[i]) + counter_advance(last0, chunk, qr))):
                ref = np.zeros((qr, chunk), np.uint16)
                for lo in (0, chunk - cols):  # the first and the last 2,048 rows of the chunk
                    ref[:, lo:lo + cols] = orc.u16_dot(rows, m.db.read_records(c0 + lo, cols), threads=threads)
                exp = orc.mask_dot_share(ref.reshape(-1), m.mult, seeds[0], seeds[1], 0, ctr).reshape(qr, chunk)
                got = got_t[:qr * chunk].cpu().numpy().view(np.uint16).reshape(qr, chunk)
                for lo in (0, chunk - cols):
                    assert np.array_equal(got[:, lo:lo + cols], exp[:, lo:lo + cols]), \
                        f"matrix {k} batch {i} ({sizes[i]} queries) chunk at {c0} columns {lo}.."
    orc.set_threads(1)


def test_c5_serving_fullsize(irismpc, orc):
    import torch

    from irismpc import serving

    _free_device(torch)
    N, U = 2_097_152, 32768
    eng = serving.ServingEngines(N, N + 2 * irismpc.DB_CHUNK_SIZE, max_active=64, seed_base=21)
    eng.build(U, 95, 96)
    dc, dm = zip(*[orc.synth_iris(95, u) for u in range(U)])  # the oracle's twin of the device generator
    dc, dm = np.array(dc), np.array(dm)
    rng = np.random.default_rng(2097152)
    plain, counter = {}, [0]

    def shared(c, m):
        counter[0] += 1
        cs, ms = orc.synth_shares(c, m, 99, counter[0])
        return [(cs[p], ms[p]) for p in range(3)]

    def uniq(k):
        if rng.random() < 0.5:  # a returning user: a DB iris with 10-30 % of its code bits flipped
            u = int(rng.integers(U))
            c, m = orc.synth_flip(dc[u], 97, k, int(rng.integers(1280, 3841))), dm[u]
        else:
            c, m = orc.synth_iris(98, k)
        r = serving.Request("uniqueness", 0, shares=shared(c, m))
        plain[id(r)] = (c, m)
        return r

    def reauth(k):
        row = int(rng.integers(N))
        u = row % U if rng.random() < 0.7 else int(rng.integers(U))  # the right iris or another one
        c, m = orc.synth_flip(dc[u], 97, 1000 + k, 1500), dm[u]
        r = serving.Request("reauth", 0, shares=shared(c, m), target=row)
        plain[id(r)] = (c, m)
        return r

    def insert(k):
        c, m = orc.synth_iris(98, 1000 + k)
        r = serving.Request("insert", 0, shares=shared(c, m))
        plain[id(r)] = (c, m)
        return r

    def delete(k):
        return serving.Request("delete", 0, target=int(rng.integers(N)))

    reqs = serving.poisson_mix(30.0, 48, rng, uniq, reauth, insert, delete, mix=(0.75, 0.15, 0.10))
    out = serving.Server(eng).run(reqs)
    deleted = {r.target for r in out if r.kind == "delete"}
    inserted = {r.result: plain[id(r)] for r in out if r.kind == "insert"}
    kinds = {r.kind for r in out}
    assert {"uniqueness", "reauth", "insert", "delete"} <= kinds
    hits = {}
    for r in out:
        if r.kind == "uniqueness":
            qc, qm = plain[id(r)]
            hit = orc.plain_match(qc, qm, dc, dm, A_0375)
            hits[id(r)] = hit
            n0 = min(N, r.db_rows)
            rows = np.flatnonzero(np.tile(hit, N // U)[:n0])
            exp = set(rows.tolist())
            for row, (ic, im) in inserted.items():
                if row < r.db_rows and orc.plain_match(qc, qm, ic[None], im[None], A_0375)[0]:
                    exp.add(row)
            assert set(r.result) - deleted == exp - deleted, f"uniqueness rid {r.rid}"
        elif r.kind == "reauth" and r.target not in deleted:
            qc, qm = plain[id(r)]
            u = r.target % U
            assert r.result == bool(orc.plain_match(qc, qm, dc[u:u + 1], dm[u:u + 1], A_0375)[0]), \
                f"reauth rid {r.rid}"
    found = [r for r in out if r.kind == "uniqueness" and r.result]
    assert len(found) >= 3, "too few returning users to exercise the matches"
    # the same requests alone (the plain per-request form, on the DB as it is now)
    for r in found[:2]:
        alone = set(eng.run_alone(r.shares))
        keep = {x for x in alone if x < r.db_rows} - deleted
        assert keep == set(r.result) - deleted, f"rid {r.rid} alone vs in the mix"
