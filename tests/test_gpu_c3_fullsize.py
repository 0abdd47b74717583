"""Full 3-party protocol at full size, both eyes: 3 parties x 2 eyes x
1,048,576 DB share rows in HBM, 64 queries x 31 rotations, threshold 0.375.

The DB holds 32,768 distinct synthetic irises per eye tiled over the rows
(row r = iris r mod 32,768) and 32 of the 64 queries are copies of DB irises
with 10-30% of their code bits flipped. The expected answer is the oracle's
plaintext brute force of every query rotation against the 32,768 distinct
irises with the integer MPC predicate, tiled over the rows (the planted
copies are among its matches). Checked: every party's opened bitmap of each
eye equals it bit for bit, and every party's AND-over-eyes merge returns the
first row matching in both eyes. Smaller multi-chunk scans are checked
against the oracle share for share in test_gpu_pipeline.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, B, U = 1 << 20, 64, 32768
A_0375 = 16384


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


def test_c3_fullsize_planted(irismpc, orc):
    import torch

    from irismpc import synth

    torch.cuda.empty_cache()  # the 6 DB shares take ~242 GB of the 288 GB
    Qr, chunk = B * 31, irismpc.DB_CHUNK_SIZE
    rng = np.random.default_rng(2024)
    targets = rng.choice(U, size=B // 2, replace=False)
    flips = rng.integers(1280, 3841, size=B // 2)
    planted = [(j, int(t), int(f)) for j, t, f in zip(range(B // 2), targets, flips)]
    base = [np.full(8, p, np.uint32) for p in range(3)]
    eyes, expects = [], []
    for eye in range(2):
        ds = 61 + eye
        ex = irismpc.Exchange(timeout_s=300)
        parties, bitmaps, streams = [], [], []
        for p in range(3):
            seed = lambda n: (irismpc.derive_seed(base[p], n), irismpc.derive_seed(base[(p + 2) % 3], n))  # noqa: E731
            codes = irismpc.ShareDB(p, chunk, Qr, 12800, seed(0), comm=ex, tag=0)
            masks = irismpc.ShareDB(p, chunk, Qr, 6400, seed(1), comm=ex, tag=1)
            circ = irismpc.Circuits(p, Qr * chunk, Qr * chunk // 64, seed(5), ex, tag=2)
            cmp = irismpc.DistanceComparator(Qr, N)
            parties.append(irismpc.PartyEngines(codes, masks, circ, cmp, codes.alloc_db(N), masks.alloc_db(N)))
            bitmaps.append(torch.zeros(N * B // 64 + 2, dtype=torch.int64, device="cuda"))
            streams.append(irismpc.stream_create())
        synth.build_db([pe.code_db for pe in parties], [pe.mask_db for pe in parties], N, U, ds, ds + 30)
        q_codes, q_masks, shares = synth.make_queries(B, planted, ds, ds + 20, ds + 10, ds + 40)
        db_c, db_m = map(np.array, zip(*[orc.synth_iris(ds, u) for u in range(U)]))
        hit = np.stack([orc.plain_match(q_codes[j], q_masks[j], db_c, db_m, A_0375) for j in range(B)])
        expects.append(np.tile(hit, (1, N // U)))
        qcs = [parties[p].codes.upload_query_shares(shares[p][0]) for p in range(3)]
        qms = [parties[p].masks.upload_query_shares(shares[p][1]) for p in range(3)]
        print(f"eye {eye}: DB built, scanning", flush=True)
        irismpc.scan_db_trio(parties, qcs, qms, N, B, A_0375, [b.data_ptr() for b in bitmaps], streams)
        for s in streams:
            irismpc.stream_synchronize(s)
        eyes.append((parties, bitmaps))

    rows = np.arange(N)
    for j, t, _ in planted:  # the planted copies are among the brute-force matches, in both eyes
        assert all(e[j][rows % U == t].all() for e in expects)
    for eye, (parties, bitmaps) in enumerate(eyes):
        expect_words = np.packbits(expects[eye].reshape(-1), bitorder="little").view(np.uint64)
        for p in range(3):
            got = bitmaps[p].cpu().numpy().view(np.uint64)[: N * B // 64]
            assert np.array_equal(got, expect_words), f"eye {eye} party {p} bitmap"
    both = expects[0] & expects[1]
    want = np.array([np.flatnonzero(both[j])[0] if both[j].any() else 0xFFFFFFFF for j in range(B)], np.uint32)
    for p in range(3):
        final = torch.full((B,), -1, dtype=torch.int32, device="cuda")
        eyes[0][0][p].comparator.reset()
        eyes[0][0][p].comparator.join_db_matches(eyes[0][1][p].data_ptr(), eyes[1][1][p].data_ptr(),
                                                 final.data_ptr(), N)
        torch.cuda.synchronize()
        assert np.array_equal(final.cpu().numpy().view(np.uint32), want), f"party {p} merged results"
