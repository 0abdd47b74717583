"""The reference's per-chunk scan steps, called one by one through the C ABI.

compare_query_against_db_and_self_prefiltering (iris-mpc-gpu/src/server/
actor.rs:2246-2589) runs, per DB chunk: dot + dot_reduce_and_multiply
(dot/share_db.rs:800-905), reshare_results (share_db.rs:1004-1036),
result_chunk_shares (share_db.rs:1050-1069), compare_threshold_masked_many
(threshold_ring/protocol.rs:2519-2546), the open exchange (actor.rs:2928-2997)
and open_results (dot/distance_comparator.rs:196-279). irismpc_scan_db_trio
fuses those steps (one joint GEMM launch for the three parties, the circuits
overlapping the next chunk). Here every step is its own call: once with the
three-party entry points from one thread, once with the per-party collective
entry points from three threads (each engine keeping a private copy of the
peer share). Both must give the bitmaps of the plaintext predicate, the scan's
circuit result shares and the scan's counters.
"""
import threading

import numpy as np
import pytest

from test_gpu_pipeline import _plain_bitmap, _shares, _workload
from test_gpu_share_db import _rand, _rotated_rows

pytestmark = pytest.mark.gpu

A_0375 = 16384
N, B, PLANTED = 2000, 8, [0, 3, 6]
QR = B * 31
WIDTH = (N + 63) // 64 * 64  # actor.rs:2331-2337: the chunk padded to 64 rows


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


@pytest.fixture(scope="module")
def workload(orc):
    import os

    orc.set_threads(min(16, os.cpu_count() or 1))
    db_c, db_m, q_c, q_m = _workload(orc, N, B, PLANTED, flips=1280)
    dbc_sh, dbm_sh = _shares(orc, db_c, db_m, 40)
    qc_sh, qm_sh = _shares(orc, q_c, q_m, 41)
    exp = _plain_bitmap(orc, db_c, db_m, q_c, q_m, A_0375)
    orc.set_threads(1)
    assert exp.sum() >= len(PLANTED)
    return dbc_sh, dbm_sh, qc_sh, qm_sh, exp


def _parties(irismpc, w):
    """Three parties' engines (seeds as the actor derives them) with the DB
    shares loaded and the queries uploaded; the exchange must outlive them."""
    import torch

    dbc_sh, dbm_sh, qc_sh, qm_sh, _ = w
    base = [np.full(8, 7 + p, np.uint32) for p in range(3)]
    seeds = [(base[p], base[(p + 2) % 3]) for p in range(3)]

    def der(p, n):
        return irismpc.derive_seed(seeds[p][0], n), irismpc.derive_seed(seeds[p][1], n)

    ex = irismpc.Exchange(timeout_s=120)
    chunk = irismpc.DB_CHUNK_SIZE
    parties, qcs, qms = [], [], []
    for p in range(3):
        codes = irismpc.ShareDB(p, chunk, QR, 12800, der(p, 0), comm=ex, tag=0)
        masks = irismpc.ShareDB(p, chunk, QR, 6400, der(p, 1), comm=ex, tag=1)
        circ = irismpc.Circuits(p, QR * chunk, QR * chunk // 64, der(p, 5), ex, tag=2)
        cmp = irismpc.DistanceComparator(QR, N)
        cdb, mdb = codes.alloc_db(N), masks.alloc_db(N)
        cdb.load_full_db(dbc_sh[p])
        mdb.load_full_db(dbm_sh[p])
        parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, cdb, mdb))
        qcs.append(codes.upload_query_shares(qc_sh[p]))
        qms.append(masks.upload_query_shares(qm_sh[p]))
    bitmaps = [torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
    streams = [irismpc.stream_create() for _ in range(3)]
    torch.cuda.synchronize()
    return ex, parties, qcs, qms, bitmaps, streams


def _bits(bitmap):
    return np.unpackbits(bitmap.cpu().numpy().view(np.uint8), bitorder="little")[:N * B].astype(bool)


def _open(irismpc, P, bitmap, s):
    """Result shares (a, b) and the opened peer view c into the bitmap."""
    a, b = P.circuits.take_result_buffer()
    va, vb, c = P.circuits.open_views()
    assert (a, b) == (va, vb)
    P.comparator.open_results(a, b, c, bitmap.data_ptr(), WIDTH, N, 0, N, B, s)
    return a, b, c


def _steps_trio(irismpc, parties, qcs, qms, bitmaps, streams):
    for P, qc, qm, s in zip(parties, qcs, qms, streams):
        P.codes.dot_reduce(qc, P.code_db, WIDTH, 0, 1, s)
        P.masks.dot_reduce(qm, P.mask_db, WIDTH, 0, 2, s)
    irismpc.ShareDB.reshare_results_trio([P.codes for P in parties], WIDTH, streams)
    irismpc.ShareDB.reshare_results_trio([P.masks for P in parties], WIDTH, streams)
    codes = [P.codes.result_chunk_shares() for P in parties]
    masks = [P.masks.result_chunk_shares() for P in parties]
    circs = [P.circuits for P in parties]
    for c in circs:
        c.set_chunk_size(WIDTH * QR // 64)
    irismpc.Circuits.compare_threshold_masked_many_trio(circs, codes, masks, WIDTH * QR, A_0375, streams)
    irismpc.Circuits.open_trio(circs, streams)
    views = [_open(irismpc, P, bm, s) for P, bm, s in zip(parties, bitmaps, streams)]
    for s in streams:
        irismpc.stream_synchronize(s)
    return views


def _steps_threads(irismpc, parties, qcs, qms, bitmaps, streams):
    views, errors = [None] * 3, []

    def party(p):
        try:
            P, s = parties[p], streams[p]
            P.codes.set_peer_copy(True)
            P.masks.set_peer_copy(True)
            P.codes.dot_reduce(qcs[p], P.code_db, WIDTH, 0, 1, s)
            P.masks.dot_reduce(qms[p], P.mask_db, WIDTH, 0, 2, s)
            P.codes.reshare_results(WIDTH, s)
            P.masks.reshare_results(WIDTH, s)
            P.circuits.set_chunk_size(WIDTH * QR // 64)
            P.circuits.compare_threshold_masked_many(P.codes.result_chunk_shares(),
                                                     P.masks.result_chunk_shares(), WIDTH * QR, A_0375, s)
            P.circuits.open(s)
            views[p] = _open(irismpc, P, bitmaps[p], s)
            irismpc.stream_synchronize(s)
        except Exception as e:  # reported by the main thread
            errors.append((p, e))

    threads = [threading.Thread(target=party, args=(p,)) for p in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=180)
    assert not any(t.is_alive() for t in threads), "a party thread did not finish"
    assert not errors, errors
    return views


@pytest.mark.parametrize("mode", ["trio", "threads"])
def test_stepwise_chunk_equals_scan(irismpc, workload, mode):
    exp = workload[4]
    ref = _parties(irismpc, workload)
    ex_r, par_r, qc_r, qm_r, bm_r, st_r = ref
    irismpc.scan_db_trio(par_r, qc_r, qm_r, N, B, A_0375, [b.data_ptr() for b in bm_r], st_r)
    for s in st_r:
        irismpc.stream_synchronize(s)

    ex, parties, qcs, qms, bitmaps, streams = _parties(irismpc, workload)
    steps = _steps_trio if mode == "trio" else _steps_threads
    views = steps(irismpc, parties, qcs, qms, bitmaps, streams)

    W = QR * WIDTH // 64
    for p in range(3):
        assert np.array_equal(_bits(bitmaps[p]), exp), f"party {p} bitmap vs plaintext"
        assert np.array_equal(_bits(bm_r[p]), exp), f"party {p} scan bitmap vs plaintext"
        ra, rb, _ = par_r[p].circuits.open_views()
        a, b, _ = views[p]
        assert np.array_equal(irismpc.download(a, W, np.uint64), irismpc.download(ra, W, np.uint64)), \
            f"party {p} result a"
        assert np.array_equal(irismpc.download(b, W, np.uint64), irismpc.download(rb, W, np.uint64)), \
            f"party {p} result b"
        P, R = parties[p], par_r[p]
        assert (P.codes.counter, P.masks.counter, P.circuits.counter) == \
            (R.codes.counter, R.masks.counter, R.circuits.counter), f"party {p} counters"

    # the same opened shares through an index map (the two-stage flow's subset
    # chunk, distance_comparator.rs:281-342): column j is DB row perm[j]
    import torch

    perm = np.random.default_rng(5).permutation(N).astype(np.uint32)
    imap = torch.from_numpy(perm.view(np.int32)).cuda()
    mapped = torch.zeros_like(bitmaps[0])
    torch.cuda.synchronize()
    P = parties[1]
    P.comparator.get_partial_results_with_rotations()  # clears the capture
    a, b, c = views[1]
    P.comparator.open_results_with_index_mapping(a, b, c, mapped.data_ptr(), WIDTH, N, N, B, imap.data_ptr())
    torch.cuda.synchronize()
    want = np.zeros((B, N), bool)
    want[:, perm] = exp.reshape(B, N)
    assert np.array_equal(_bits(mapped).reshape(B, N), want)
    rot = P.comparator.get_partial_results_with_rotations()
    got = {(q, d) for q, ds in rot.items() for d in ds}
    assert got == set(zip(*np.nonzero(want)))
    assert all(-15 <= r <= 15 for ds in rot.values() for rs in ds.values() for r in rs)


@pytest.mark.parametrize("K,mult", [(12800, 1), (6400, 2)])
def test_dot_reduce_packed_equals_per_batch_ranges(irismpc, orc, K, mult):
    """Batching form: three batches (2, 3 and 1 queries) packed in one query set
    and one launch give each batch the shares (values, layout, zero-share
    counters) of its own dot_reduce_range call; the first chunk of the first
    batch against the oracle."""
    import torch

    rng = np.random.default_rng(21 + mult)
    sizes, Nd, chunk = [2, 3, 1], 640, 256  # chunks of 256, 256 and 128 DB rows
    raws = [_rand(rng, (b, K)) for b in sizes]
    rows = [31 * b for b in sizes]
    db = _rand(rng, (Nd, K))
    seeds = [np.full(8, 11, np.uint32), np.full(8, 12, np.uint32)]
    ex = irismpc.Exchange()
    packed = irismpc.ShareDB(1, chunk, sum(rows), K, seeds, comm=ex, tag=0)
    single = irismpc.ShareDB(1, chunk, sum(rows), K, seeds, comm=ex, tag=1)
    dbp, dbs = packed.alloc_db(Nd), single.alloc_db(Nd)
    dbp.load_full_db(db)
    dbs.load_full_db(db)
    qp = packed.upload_query_shares_rows(np.concatenate(raws), sum(rows))
    outs_p = [torch.zeros(r * Nd + 64, dtype=torch.int16, device="cuda") for r in rows]
    outs_s = [torch.zeros(r * Nd + 64, dtype=torch.int16, device="cuda") for r in rows]
    torch.cuda.synchronize()
    packed.dot_reduce_packed(qp, dbp, 0, Nd, chunk, mult, rows, [o.data_ptr() for o in outs_p])
    for raw, r, o in zip(raws, rows, outs_s):
        q = single.upload_query_shares_rows(raw, r)
        single.dot_reduce_range(q, dbs, 0, Nd, chunk, mult, o.data_ptr())
    torch.cuda.synchronize()
    assert packed.counter == single.counter > 0
    for i, (op, os_) in enumerate(zip(outs_p, outs_s)):
        assert np.array_equal(op.cpu().numpy(), os_.cpu().numpy()), f"batch {i}"
    ref = orc.u16_dot(_rotated_rows(orc, 1, raws[0], rows[0]), db[:chunk]).reshape(-1)
    ref = orc.mask_dot_share(ref, mult, seeds[0], seeds[1], 0, 0)
    assert np.array_equal(outs_p[0].cpu().numpy().view(np.uint16)[:rows[0] * chunk], ref)


def test_device_memory_helpers(irismpc):
    """The device-memory entry points a host without torch uses (irismpc_malloc,
    irismpc_fill_u32, irismpc_fill_u64, irismpc_memcpy_dtoh, irismpc_free),
    with irismpc_version and irismpc_device_count."""
    import ctypes

    from irismpc._lib import call, lib

    assert lib().irismpc_version() > 0
    n = ctypes.c_int()
    call("irismpc_device_count", ctypes.byref(n))
    assert n.value >= 1
    p = ctypes.c_void_p()
    call("irismpc_malloc", 0, 4096, ctypes.byref(p))
    assert p.value
    try:
        call("irismpc_fill_u32", p.value, 1024, 0xDEADBEEF, None)
        call("irismpc_fill_u64", p.value + 1024, 64, 0x0123456789ABCDEF, None)
        got = irismpc.download(p.value, 1024, np.uint32)
        want = np.full(1024, 0xDEADBEEF, np.uint32)
        want[256:384] = np.full(64, 0x0123456789ABCDEF, np.uint64).view(np.uint32)
        assert np.array_equal(got, want)
    finally:
        call("irismpc_free", p.value)
