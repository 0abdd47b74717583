"""GPU parity of the share GEMM (ShareDB) against the oracle.

check_matmul analogue (iris-mpc-gpu/src/dot/share_db.rs:1133-1198): random u16
matrices, every GEMM tile variant, code (K=12800, x1) and trimmed mask (K=6400,
x2) shapes, plus the masked multi-chunk form with the reference ChaCha12
counter schedule.
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


def _rand(rng, shape):
    return rng.integers(0, 1 << 16, shape, dtype=np.uint16)


@pytest.mark.parametrize("K,mult", [(12800, 1), (6400, 2)])
def test_share_gemm_all_variants(irismpc, orc, K, mult):
    rng = np.random.default_rng(42)
    Qr, N = 248, 1000
    chunk = 1024  # padded to 64 like the reference's dot_chunk_size
    q = _rand(rng, (Qr, K))
    db = _rand(rng, (N, K))
    ref = orc.u16_dot(q, db)
    ref = (ref.astype(np.uint32) * mult).astype(np.uint16)
    eng = irismpc.ShareDB(0, chunk, Qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    assert np.array_equal(slices.read_records(0, N), db)
    dq = eng.upload_query(q)
    for v in range(irismpc.gemm_variants()):
        eng.set_variant(v)
        eng.dot_reduce(dq, slices, chunk, 0, mult)
        got = eng.fetch_results(Qr * chunk).reshape(Qr, chunk)
        assert np.array_equal(got[:, :N], ref), f"variant {v}"


def test_share_gemm_offset_and_limb_load(irismpc, orc):
    """DB loaded from S3-style split limbs; chunk at a non-zero offset."""
    rng = np.random.default_rng(3)
    K, Qr, N = 6400, 248, 700
    q = _rand(rng, (Qr, K))
    db = _rand(rng, (N, K))
    l0, l1 = orc.limbs(db)
    eng = irismpc.ShareDB(1, 512, Qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_single_record_from_s3(0, l0.view(np.uint8), l1.view(np.uint8))
    assert np.array_equal(slices.read_records(0, N), db)
    dq = eng.upload_query(q)
    eng.dot_reduce(dq, slices, 192, 512, 1)  # rows 512..703 (padded chunk 192)
    got = eng.fetch_results(Qr * 192).reshape(Qr, 192)
    ref = orc.u16_dot(q, db[512:700])
    assert np.array_equal(got[:, :188], ref)


def test_share_gemm_masked_chunks(irismpc, orc):
    """Remote engine: zero-share masks from the ChaCha12 streams at the
    reference counters, multi-chunk single launch vs per-chunk launches."""
    import torch

    rng = np.random.default_rng(4)
    K, Qr, N = 12800, 248, 640
    chunk = 256
    q = _rand(rng, (Qr, K))
    db = _rand(rng, (N, K))
    seeds = [np.full(8, 7, np.uint32), np.full(8, 9, np.uint32)]
    ex = irismpc.Exchange()
    eng = irismpc.ShareDB(0, chunk, Qr, K, seeds, comm=ex, tag=0)
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query(q)
    eng.counter = 1000
    out = torch.zeros(N * Qr + 64, dtype=torch.int16, device="cuda")
    eng.dot_reduce_range(dq, slices, 0, N, chunk, 1, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)
    ctr = 1000
    for c0 in range(0, N, chunk):
        w = min(chunk, N - c0)
        ref = orc.u16_dot(q, db[c0:c0 + w]).reshape(-1)
        ref = orc.mask_dot_share(ref, 1, seeds[0], seeds[1], 0, ctr)
        assert np.array_equal(got[c0 * Qr:c0 * Qr + w * Qr], ref), f"chunk at {c0}"
        ctr += 2 * ((w * Qr * 2 + 63) // 64)
    assert eng.counter == ctr
    # per-chunk launch of the first chunk at the same counter
    eng.counter = 1000
    eng.dot_reduce(dq, slices, chunk, 0, 1)
    got1 = eng.fetch_results(Qr * chunk)
    assert np.array_equal(got1, got[:Qr * chunk])


def _rotated_rows(orc, party, raw, qr):
    """The reference's query preparation: Lagrange/basis B per share, then the
    31 rotations of every query; rows past the batch are zero shares."""
    rows = np.zeros((qr, raw.shape[1]), np.uint16)
    for j in range(raw.shape[0]):
        rows[j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(party, raw[j]))
    return rows


@pytest.mark.parametrize("K,mult,batch,qr", [(12800, 1, 6, 250), (6400, 2, 9, 279), (12800, 1, 1, 31),
                                              (6400, 2, 2, 62), (12800, 1, 1, 40), (12800, 1, 5, 155),
                                              (6400, 2, 10, 310)])
def test_share_gemm_rotation_variants(irismpc, orc, K, mult, batch, qr):
    """Rotation-aware kernels on queries uploaded from raw shares (device
    preprocessing + rotations) against the oracle's rotated rows, next to the
    generic variants and the automatic choice; query_length is not a multiple
    of the 4-query tile, so the last tile is partial; one and two queries
    (where the automatic choice is the two-query tile), also with spare rows;
    5, 9 and 10 queries (where it takes the two-query tile too)."""
    rng = np.random.default_rng(11)
    N, chunk, party = 1000, 1024, 1
    raw = _rand(rng, (batch, K))
    db = _rand(rng, (N, K))
    ref = orc.u16_dot(_rotated_rows(orc, party, raw, qr), db)
    ref = (ref.astype(np.uint32) * mult).astype(np.uint16)
    eng = irismpc.ShareDB(party, chunk, qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares(raw)
    for v in [-1] + list(range(irismpc.gemm_variants())) + irismpc.gemm_rot_variants():
        eng.set_variant(v)
        eng.dot_reduce(dq, slices, chunk, 0, mult)
        got = eng.fetch_results(qr * chunk).reshape(qr, chunk)
        assert np.array_equal(got[:, :N], ref), f"variant {v}"


@pytest.mark.parametrize("K,mult", [(12800, 1), (6400, 2)])
def test_query_refill_matches_fresh_upload(irismpc, orc, K, mult):
    """A query set refilled in place with other shares and fewer queries (as
    the request loop does when its active set changes) gives the dot shares of
    a fresh upload of those shares, under the automatic and the generic GEMM
    variants; its DB form follows the new shares; a refill larger than the set
    is rejected."""
    rng = np.random.default_rng(12)
    N, chunk, party = 1000, 1024, 0
    first, second = _rand(rng, (8, K)), _rand(rng, (5, K))
    db = _rand(rng, (N, K))
    eng = irismpc.ShareDB(party, chunk, 8 * 31, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares_rows(first, 8 * 31)
    for raw in (second, first[:3], second):
        qr = raw.shape[0] * 31
        ref = orc.u16_dot(_rotated_rows(orc, party, raw, qr), db)
        ref = (ref.astype(np.uint32) * mult).astype(np.uint16)
        eng.refill_query_shares(dq, raw, qr)
        for v in (-1, 0):
            eng.set_variant(v)
            eng.dot_reduce(dq, slices, chunk, 0, mult)
            got = eng.fetch_results(qr * chunk).reshape(qr, chunk)
            assert np.array_equal(got[:, :N], ref), f"variant {v}, batch {raw.shape[0]}"
        eng.set_variant(-1)
        slices.write_at_index(dq, raw.shape[0] - 1, 11)
        assert np.array_equal(slices.read_records(11, 1)[0], raw[-1])
        db[11] = raw[-1]  # the DB now holds that row
    with pytest.raises(RuntimeError):
        eng.refill_query_shares(dq, _rand(rng, (9, K)), 9 * 31)


def test_query_refill_async_is_stream_ordered(irismpc, orc):
    """Asynchronous refills queued on one stream between GEMMs that read the
    set: each GEMM sees the shares of the refill queued before it (from pinned
    host memory, no synchronisation in between), equal to the oracle's dots."""
    import torch

    rng = np.random.default_rng(13)
    K, N, chunk = 6400, 700, 768
    raws = [_rand(rng, (b, K)) for b in (4, 2, 3)]
    db = _rand(rng, (N, K))
    eng = irismpc.ShareDB(1, chunk, 4 * 31, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares_rows(raws[0], 4 * 31)
    stream = irismpc.stream_create()
    pinned = [torch.empty(r.shape, dtype=torch.int16, pin_memory=True) for r in raws]
    outs = [torch.zeros(4 * 31 * chunk + 64, dtype=torch.int16, device="cuda") for _ in raws]
    torch.cuda.synchronize()
    for raw, pin, out in zip(raws, pinned, outs):
        pin.numpy().view(np.uint16)[:] = raw
        eng.refill_query_shares_async(dq, pin.data_ptr(), raw.shape[0], raw.shape[0] * 31, stream)
        eng.dot_reduce_range(dq, slices, 0, chunk, chunk, 2, out.data_ptr(), stream)
    irismpc.stream_synchronize(stream)
    for raw, out in zip(raws, outs):
        qr = raw.shape[0] * 31
        ref = (orc.u16_dot(_rotated_rows(orc, 1, raw, qr), db).astype(np.uint32) * 2).astype(np.uint16)
        got = out[:qr * chunk].cpu().numpy().view(np.uint16).reshape(qr, chunk)
        assert np.array_equal(got[:, :N], ref), f"batch of {raw.shape[0]}"


def test_share_gemm_rotation_masked_range(irismpc, orc):
    """Masked multi-chunk launch with every rotation-aware kernel (4-query
    tile and the one- and two-query tiles): same shares as the generic kernel,
    first chunk checked against the oracle."""
    import torch

    rng = np.random.default_rng(12)
    K, qr, batch, N, chunk = 12800, 248, 8, 768, 256
    raw = _rand(rng, (batch, K))
    db = _rand(rng, (N, K))
    seeds = [np.full(8, 3, np.uint32), np.full(8, 5, np.uint32)]
    ex = irismpc.Exchange()
    eng = irismpc.ShareDB(2, chunk, qr, K, seeds, comm=ex, tag=0)
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares(raw)
    outs = []
    variants = irismpc.gemm_rot_variants() + [0]
    for v in variants:
        eng.set_variant(v)
        eng.counter = 40
        out = torch.zeros(N * qr + 64, dtype=torch.int16, device="cuda")
        eng.dot_reduce_range(dq, slices, 0, N, chunk, 1, out.data_ptr())
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy().view(np.uint16))
    for v, o in zip(variants[1:], outs[1:]):
        assert np.array_equal(outs[0], o), f"variant {v}"
    ref = orc.u16_dot(_rotated_rows(orc, 2, raw, qr), db[:chunk]).reshape(-1)
    ref = orc.mask_dot_share(ref, 1, seeds[0], seeds[1], 0, 40)
    assert np.array_equal(outs[0][:chunk * qr], ref)


@pytest.mark.parametrize("K,mult,batch", [(12800, 1, 1), (6400, 2, 3)])
def test_dot_reduce_chunks_equals_per_chunk_calls(irismpc, orc, K, mult, batch):
    """One launch over consecutive chunks (two full, one partial) gives the
    shares, masks included, of per-chunk dot_reduce calls made in order, and
    leaves the engine counter where they leave it; the partial chunk checked
    against the oracle at its counter."""
    rng = np.random.default_rng(21)
    chunk, total, N = 256, 640, 768
    qr = batch * 31
    raw = _rand(rng, (batch, K))
    db = _rand(rng, (N, K))
    seeds = [np.full(8, 7, np.uint32), np.full(8, 9, np.uint32)]
    ex = irismpc.Exchange()
    eng = irismpc.ShareDB(2, total, qr, K, seeds, comm=ex, tag=0)
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares(raw)
    eng.counter = 40
    parts, ctrs = [], []
    for c0 in range(0, total, chunk):
        w = min(chunk, total - c0)
        ctrs.append(eng.counter)
        eng.dot_reduce(dq, slices, w, c0, mult)
        parts.append(eng.fetch_results(qr * w))
    per_chunk, ctr_end = np.concatenate(parts), eng.counter
    eng.counter = 40
    eng.dot_reduce_chunks(dq, slices, 0, total, chunk, mult)
    got = eng.fetch_results(qr * total)
    assert np.array_equal(got, per_chunk)
    assert eng.counter == ctr_end
    last = orc.u16_dot(_rotated_rows(orc, 2, raw, qr), db[512:640]).reshape(-1)
    last = orc.mask_dot_share(last, mult, seeds[0], seeds[1], 0, ctrs[-1])
    assert np.array_equal(got[512 * qr:], last)


def test_dot_reduce_chunks_rejects_bad_geometry(irismpc):
    """Chunk runs that do not fit the result buffer or the DB, or whose sizes
    break the tiling, fail with an error before any launch."""
    K, qr = 6400, 31
    eng = irismpc.ShareDB(0, 512, qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(512)
    dq = eng.upload_query_shares(_rand(np.random.default_rng(2), (1, K)))
    for offset, total, chunk in ((0, 100, 256), (0, 512, 100), (0, 1024, 256), (8, 256, 256),
                                 (1024, 256, 256)):
        with pytest.raises(irismpc.IrisMpcError):
            eng.dot_reduce_chunks(dq, slices, offset, total, chunk, 1)
    eng.dot_reduce_chunks(dq, slices, 0, 512, 256, 1)  # a fitting run still works
    with pytest.raises(irismpc.IrisMpcError):
        eng.dot_reduce_chunks(dq, slices, 0, 256, 256, 1, out_offset=512 * qr)


def test_rotation_variant_needs_structured_query(irismpc):
    eng = irismpc.ShareDB(0, 256, 248, 6400, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(256)
    dq = eng.upload_query(np.zeros((248, 6400), np.uint16))
    eng.set_variant(irismpc.gemm_rot_variants()[0])
    with pytest.raises(irismpc.IrisMpcError, match="raw shares"):
        eng.dot_reduce(dq, slices, 256, 0, 1)


def test_write_at_index(irismpc, orc):
    rng = np.random.default_rng(5)
    K, Qr = 6400, 248
    q = _rand(rng, (Qr, K))
    eng = irismpc.ShareDB(0, 256, Qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(300)
    slices.load_full_db(_rand(rng, (300, K)))
    dq = eng.upload_query(q)
    slices.write_at_index(dq, 3, 17)
    # write_at_index copies the centre rotation (row 3*31 + 15) of the query batch
    assert np.array_equal(slices.read_records(17, 1)[0], q[3 * 31 + 15])


def test_write_at_index_insert_form(irismpc, orc):
    """Query sets uploaded from raw shares carry the DB form of each query
    (the reference's code_query_insert centre rotation): write_at_index stores
    the raw share itself, not the Lagrange-interpolated query row."""
    rng = np.random.default_rng(6)
    K, B = 12800, 5
    raw = _rand(rng, (B, K))
    eng = irismpc.ShareDB(1, 256, B * 31, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(64)
    slices.load_full_db(_rand(rng, (64, K)))
    dq = eng.upload_query_shares(raw)
    slices.write_at_index(dq, 3, 40)
    assert np.array_equal(slices.read_records(40, 1)[0], raw[3])
    with pytest.raises(irismpc.IrisMpcError, match="beyond query batch"):
        slices.write_at_index(dq, 5, 41)


@pytest.mark.parametrize("K,flip", [(12800, True), (6400, False)])
def test_mirrored_query_upload(irismpc, orc, K, flip):
    """Mirror-orientation queries: rows = 31 rotations of the preprocessed
    mirrored share (mirrored_code with the imaginary half negated for codes,
    mirrored_mask for masks), checked through the GEMM against the oracle; the
    DB form is the separately given shares."""
    rng = np.random.default_rng(13)
    N, chunk, party, batch = 300, 512, 2, 3
    qr = batch * 31
    raw, own = _rand(rng, (batch, K)), _rand(rng, (batch, K))
    db = _rand(rng, (N, K))
    rows = np.zeros((qr, K), np.uint16)
    for j in range(batch):
        rows[j * 31:(j + 1) * 31] = orc.share_rotations(
            orc.preprocess_query(party, orc.share_mirrored(raw[j], flip)))
    ref = orc.u16_dot(rows, db)
    eng = irismpc.ShareDB(party, chunk, qr, K, ([0] * 8, [0] * 8))
    slices = eng.alloc_db(N)
    slices.load_full_db(db)
    dq = eng.upload_query_shares_mirrored(raw, own)
    eng.dot_reduce(dq, slices, chunk, 0, 1)
    got = eng.fetch_results(qr * chunk).reshape(qr, chunk)
    assert np.array_equal(got[:, :N], ref)
    slices.write_at_index(dq, 1, 7)
    assert np.array_equal(slices.read_records(7, 1)[0], own[1])
