"""Synthetic workload generator (csrc/synth.hip) against the oracle's twin
(oracle/src/synth.c): the same ChaCha20 streams must give identical plaintext
irises, bit flips and shares. The host functions run without a GPU."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "iris-mpc_amd", "irismpc", "libirismpc.so")


@pytest.fixture(scope="module")
def synth():
    if not os.path.exists(LIB):
        pytest.skip("libirismpc.so not built")
    from irismpc import synth as s

    return s


@pytest.mark.parametrize("ds,idx", [(0, 0), (1, 17), (30, 123456)])
def test_host_plaintext_matches_oracle(synth, orc, ds, idx):
    c, m = synth.synth_iris(ds, idx)
    oc, om = orc.synth_iris(ds, idx)
    assert np.array_equal(c, oc) and np.array_equal(m, om)
    # mask bits come in pairs (2k, 2k + 1) with density ~0.85
    bits = np.unpackbits(m.view(np.uint8), bitorder="little")
    assert np.array_equal(bits[0::2], bits[1::2])
    assert 0.8 < bits.mean() < 0.9


@pytest.mark.parametrize("n", [0, 1, 1920, 3840, 12800])
def test_host_flips_match_oracle(synth, orc, n):
    c, _ = synth.synth_iris(3, 9)
    got = synth.synth_flip(c, 11, 5, n)
    assert np.array_equal(got, orc.synth_flip(c, 11, 5, n))
    assert int(np.bitwise_count(got ^ c).sum()) == n


@pytest.mark.gpu
def test_device_plaintext_and_shares_match_oracle(synth, orc):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    codes, masks = synth.plaintext_device(7, 1000, 5)
    torch.cuda.synchronize()
    codes, masks = codes.cpu().numpy().view(np.uint64), masks.cpu().numpy().view(np.uint64)
    for k in range(5):
        oc, om = orc.synth_iris(7, 1000 + k)
        assert np.array_equal(codes[k], oc) and np.array_equal(masks[k], om)
    shares = synth.shares_of(codes, masks, 30, 1000)
    for k in range(5):
        a, b = orc.synth_shares(codes[k], masks[k], 30, 1000 + k)
        for p in range(3):
            assert np.array_equal(shares[p][0][k], a[p]), f"code share party {p} iris {k}"
            assert np.array_equal(shares[p][1][k], b[p]), f"mask share party {p} iris {k}"


@pytest.mark.gpu
def test_build_db_tiles_shares(synth, orc):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc

    N, U = 1000, 64
    eng = [irismpc.ShareDB(p, 1024, 31, 12800, ([0] * 8, [0] * 8)) for p in range(3)]
    meng = [irismpc.ShareDB(p, 1024, 31, 6400, ([0] * 8, [0] * 8)) for p in range(3)]
    cdb = [e.alloc_db(N) for e in eng]
    mdb = [e.alloc_db(N) for e in meng]
    synth.build_db(cdb, mdb, N, U, 5, 50)
    for r in (0, 63, 64, 511, 960, 999):
        c, m = orc.synth_iris(5, r % U)
        a, b = orc.synth_shares(c, m, 50, r % U)
        for p in range(3):
            assert np.array_equal(cdb[p].read_records(r, 1)[0], a[p]), f"row {r} party {p}"
            assert np.array_equal(mdb[p].read_records(r, 1)[0], b[p]), f"row {r} party {p}"
