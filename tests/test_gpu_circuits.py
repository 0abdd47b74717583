"""GPU parity of the threshold circuits (Circuits) against the oracle.

tests/threshold.rs analogue: random replicated shares of sampled code/mask dots,
the three parties on one GPU; the result bit shares must equal the oracle's
share for share (same ChaCha12 schedule), and the opened bits the closed form
msb(M*A - (c<<16) - 1).
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS_OWN = np.array([[0] * 8, [1] * 8, [2] * 8], np.uint32)
SEEDS_PREV = np.array([[2] * 8, [0] * 8, [1] * 8], np.uint32)


@pytest.fixture(scope="module")
def irismpc():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc as m

    m.lib()
    return m


def _sample(rng, L):
    x = rng.integers(0, 12801, L).astype(np.int64)
    neg = rng.integers(0, 2, L).astype(bool)
    code = np.where(neg, (-x) & 0xFFFF, x).astype(np.uint16)
    mask = rng.integers(0, 12801, L).astype(np.uint16)
    return code, mask


def _rep(rng, vals):
    a = rng.integers(0, 1 << 16, vals.size, dtype=np.uint16)
    b = rng.integers(0, 1 << 16, vals.size, dtype=np.uint16)
    c = (vals.astype(np.int64) - a - b).astype(np.uint16)
    x = [a, b, c]
    return [x[p] for p in range(3)], [x[(p + 2) % 3] for p in range(3)]


def _dev(torch, arr):
    return torch.from_numpy(arr.view(np.int16).copy()).cuda()


def _setup(irismpc, torch, W, alloc, counter=0, seeds=(SEEDS_OWN, SEEDS_PREV)):
    ex = irismpc.Exchange()
    circ = [irismpc.Circuits(p, 64 * W, alloc, (seeds[0][p], seeds[1][p]), ex) for p in range(3)]
    for c in circ:
        c.counter = counter
    streams = [irismpc.stream_create() for _ in range(3)]
    return ex, circ, streams


def _run_case(irismpc, orc, torch, W, alloc, A, seed, counter=0, threads=False, sampler=None,
              seeds=(SEEDS_OWN, SEEDS_PREV), alias_b=False):
    """alias_b: every party's b views are the previous party's a buffers (as the
    scan's zero-copy reshare hands them over), which selects the MSB adder form
    that reads the b planes from the previous party's wave; otherwise the b
    shares are separate buffers of the same values."""
    rng = np.random.default_rng(seed)
    L = 64 * W
    code, mask = (sampler or _sample)(rng, L)
    ca, cb = _rep(rng, code)
    ma, mb = _rep(rng, mask)
    ex, circ, streams = _setup(irismpc, torch, W, alloc, counter, seeds)
    dca = [_dev(torch, x) for x in ca]
    dcb = [_dev(torch, x) for x in cb]
    dma = [_dev(torch, x) for x in ma]
    dmb = [_dev(torch, x) for x in mb]
    if alias_b:
        dcb = [dca[(p + 2) % 3] for p in range(3)]
        dmb = [dma[(p + 2) % 3] for p in range(3)]
    lifted = [(torch.zeros(L, dtype=torch.int32, device="cuda"),
               torch.zeros(L, dtype=torch.int32, device="cuda")) for _ in range(3)]
    for p in range(3):
        circ[p].set_lift_capture(lifted[p][0].data_ptr(), lifted[p][1].data_ptr())
    torch.cuda.synchronize()
    codes = [(dca[p].data_ptr(), dcb[p].data_ptr()) for p in range(3)]
    masks = [(dma[p].data_ptr(), dmb[p].data_ptr()) for p in range(3)]
    if threads:
        errs = []

        def party(p):
            try:
                circ[p].compare_threshold_masked_many(codes[p], masks[p], L, A, streams[p])
                circ[p].open(streams[p])
            except Exception as e:  # noqa: BLE001
                errs.append(e)

        ts = [threading.Thread(target=party, args=(p,)) for p in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errs, errs
    else:
        irismpc.Circuits.compare_threshold_masked_many_trio(circ, codes, masks, L, A, streams)
        irismpc.Circuits.open_trio(circ, streams)
    for s in streams:
        irismpc.stream_synchronize(s)
    def fetch(ptr, n):
        return irismpc.download(ptr, n, np.uint64)

    views = [circ[p].open_views() for p in range(3)]
    got_a = [fetch(v[0], W) for v in views]
    got_b = [fetch(v[1], W) for v in views]
    got_c = [fetch(v[2], W) for v in views]
    ra, rb, ctr, lift_ref = orc.threshold(ca, cb, ma, mb, alloc, A, seeds[0], seeds[1],
                                          nonce=0, counter=counter, want_lifted=True)
    for p in range(3):
        assert np.array_equal(lifted[p][0].cpu().numpy().view(np.uint32), lift_ref[p]), f"lift p{p}"
        assert np.array_equal(lifted[p][1].cpu().numpy().view(np.uint32), lift_ref[(p + 2) % 3]), \
            f"lift b p{p}"
        assert np.array_equal(got_a[p], ra[p]), f"party {p} share a"
        assert np.array_equal(got_b[p], rb[p]), f"party {p} share b"
        assert np.array_equal(got_c[p], rb[(p + 2) % 3]), f"party {p} opened c"
        assert circ[p].counter == ctr
    opened = got_a[0] ^ got_b[0] ^ got_c[0]
    bits = np.unpackbits(opened.view(np.uint8), bitorder="little").astype(bool)
    exp = orc.mpc_predicate_many(code, mask, A)
    assert np.array_equal(bits, exp)
    for p in range(3):
        o = views[p]
        mine = fetch(o[0], W) ^ fetch(o[1], W) ^ fetch(o[2], W)
        assert np.array_equal(mine, opened), "all parties open the same bits"


@pytest.mark.parametrize("alias_b", [False, True])
@pytest.mark.parametrize("A", [20316, 16384])
def test_threshold_share_parity(irismpc, orc, A, alias_b):
    import torch

    _run_case(irismpc, orc, torch, W=512, alloc=512, A=A, seed=0xDEADBEEF, alias_b=alias_b)


def test_threshold_alloc_and_counter(irismpc, orc):
    """alloc > chunk (bit-inject draws cover the alloc) and a non-zero counter."""
    import torch

    _run_case(irismpc, orc, torch, W=256, alloc=1024, A=16384, seed=11, counter=123456)


def test_threshold_three_party_threads(irismpc, orc):
    """Per-party collective calls from three threads (like the reference's
    three spawn_blocking actors in tests/threshold.rs)."""
    import torch

    _run_case(irismpc, orc, torch, W=256, alloc=256, A=20316, seed=12, threads=True)


@pytest.mark.parametrize("alias_b", [False, True])
@pytest.mark.parametrize("W,alloc", [(264, 264), (40, 72)])
def test_threshold_partial_blocks(irismpc, orc, W, alloc, alias_b):
    """Word counts that leave the last block of the word adders (32 words) and
    of the bit injection (2,048 elements) partly empty, with the allocation
    larger than the chunk in the second case."""
    import torch

    _run_case(irismpc, orc, torch, W=W, alloc=alloc, A=16384, seed=W, alias_b=alias_b)


def _boundary_sample(A):
    """Code/mask dots at the threshold: c = ceil(M A / 2^16) + d, d in
    {-2, -1, 0, 1}, where msb(M A - (c << 16) - 1) flips between d = -1 and
    d = 0 (the CloseToThreshold cases of the reference's e2e generator,
    iris-mpc-common/src/test.rs:1181-1220, at the integer predicate)."""
    def sample(rng, L):
        mask = rng.integers(0, 12801, L).astype(np.int64)
        d = rng.integers(-2, 2, L)
        c = np.clip(-((-mask * A) // 65536) + d, -12800, 12800)
        return (c & 0xFFFF).astype(np.uint16), mask.astype(np.uint16)
    return sample


@pytest.mark.parametrize("alias_b", [False, True])
@pytest.mark.parametrize("A", [20316, 16384])
def test_threshold_boundary(irismpc, orc, A, alias_b):
    """Every element sits within two of the threshold; shares and opened bits
    must equal the oracle's, so a lift or comparison off by one shows."""
    import torch

    _run_case(irismpc, orc, torch, W=256, alloc=256, A=A, seed=A, sampler=_boundary_sample(A),
              alias_b=alias_b)


def test_threshold_unpaired_seeds_rejected(irismpc):
    """Seeds that are not the trio's pairwise seeds (party p's prev seed is not
    party p - 1's own seed) are rejected before any launch: the exchange
    between the parties' waves skips the reference's OTP, which is the identity
    only for pairwise seeds (the oracle, which applies the OTP, would differ)."""
    import torch

    own = np.array([[10 + p] * 8 for p in range(3)], np.uint32)
    prev = np.array([[20 + p] * 8 for p in range(3)], np.uint32)
    W = 64
    ex, circ, streams = _setup(irismpc, torch, W, W, 0, (own, prev))
    buf = torch.zeros(64 * W, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    ptrs = [(buf.data_ptr(), buf.data_ptr())] * 3
    with pytest.raises(RuntimeError, match="pairwise"):
        irismpc.Circuits.compare_threshold_masked_many_trio(circ, ptrs, ptrs, 64 * W, 16384, streams)
    assert all(c.counter == 0 for c in circ)


@pytest.mark.parametrize("alias_b", [False, True])
def test_threshold_full_chunk(irismpc, orc, alias_b):
    """One full chunk of the scan: 1,984 query-rotation rows x 32,768 DB rows =
    1,015,808 result words, share for share against the oracle (its simulator
    on the host's threads)."""
    import os

    import torch

    orc.set_threads(min(16, os.cpu_count() or 1))
    try:
        W = 1984 * 32768 // 64
        _run_case(irismpc, orc, torch, W=W, alloc=W, A=16384, seed=77, alias_b=alias_b)
    finally:
        orc.set_threads(1)
