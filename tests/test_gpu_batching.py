"""Batching form: packed launches against plain per-batch launches.

A packed launch serves several query batches of uneven size at once; every
batch must come out with exactly the shares of its own plain run (same
result layout, same ChaCha12 mask counters), also when the packed schedule
is replayed from a HIP graph, and the plain run itself must equal the
oracle's rotated-row GEMM for a sample batch.
"""
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


@pytest.mark.parametrize("graph,persistent", [(False, False), (True, False), (False, True), (True, True)])
def test_packed_equals_plain(irismpc, orc, graph, persistent):
    import torch

    from irismpc.batching import BatchRunner, Matrix

    sizes = [1, 3, 6, 2, 9, 1, 5]
    N, chunk = 70000, 32768  # three chunks, the last one partial
    rng = np.random.default_rng(9)
    ex = irismpc.Exchange()
    seeds = (np.full(8, 4, np.uint32), np.full(8, 6, np.uint32))
    mats = []
    for tag, (K, mult) in enumerate(((12800, 1), (6400, 2))):
        eng = irismpc.ShareDB(1, chunk, 256 * 31, K, seeds, comm=ex, tag=tag)
        db = eng.alloc_db(N)
        db.fill_random(0, N, np.arange(8) + 3 * tag, 0)
        shares = [rng.integers(0, 1 << 16, (b, K), dtype=np.uint16) for b in sizes]
        mats.append(Matrix(eng, db, N, mult, shares))
    stream = irismpc.stream_create()
    r = BatchRunner(mats, sizes, chunk, stream, max_queries=16)
    assert len(r.plan) > 1 and any(len(p) > 1 for p in r.plan)
    r.run_plain(ctr0=500)
    plain = [[o.clone() for o in outs] for outs in r.outs]
    ctr_after = [m.engine.counter for m in mats]
    for outs in r.outs:
        for o in outs:
            o.zero_()
    torch.cuda.synchronize()  # zeroed on torch's stream before the library's stream writes
    r.run_packed(ctr0=500, graph=graph, persistent=persistent)
    if graph:  # a second replay gives the same shares again
        for outs in r.outs:
            for o in outs:
                o.zero_()
        torch.cuda.synchronize()
        r.run_packed(ctr0=500, graph=True, persistent=persistent)
    torch.cuda.synchronize()
    for k in range(2):
        for i in range(len(sizes)):
            n = chunk * sizes[i] * 31  # the result region (the buffers carry 64 elements of slack)
            a, b = r.outs[k][i][:n].cpu().numpy(), plain[k][i][:n].cpu().numpy()
            d = np.nonzero(a != b)[0]
            assert len(d) == 0, (f"matrix {k} batch {i}: {len(d)} of {a.size} differ, first {d[:4]}, "
                                 f"last {d[-4:]}, zeros {int(np.sum(a[d] == 0))}")
    assert [m.engine.counter for m in mats] == ctr_after  # explicit counters leave the engine alone

    # the plain run's last chunk for one batch against the oracle (masks included)
    i, k = 2, 0
    m = mats[k]
    rows = np.zeros((sizes[i] * 31, 12800), np.uint16)
    for j in range(sizes[i]):
        rows[j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(1, m.shares[i][j]))
    last0, w = 65536, N - 65536
    dbrows = m.db.read_records(last0, w)
    ref = orc.u16_dot(rows, dbrows)
    ref_pad = np.zeros((rows.shape[0], 4480), np.uint16)  # the launch covers 64-padded rows
    ref_pad[:, :w] = ref
    from irismpc.batching import counter_advance, range_advance

    ctr = 500 + sum(range_advance(N, chunk, b * 31) for b in sizes[:i]) + counter_advance(last0, chunk, sizes[i] * 31)
    exp = orc.mask_dot_share(ref_pad.reshape(-1), 1, seeds[0], seeds[1], 0, ctr).reshape(rows.shape[0], 4480)
    got = plain[k][i][: rows.shape[0] * 4480].cpu().numpy().view(np.uint16).reshape(rows.shape[0], 4480)
    assert np.array_equal(got[:, :w], exp[:, :w])
