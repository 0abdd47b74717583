"""DB-chunk record format (iris-mpc-store/src/s3_importer.rs:13-94,
iris-mpc-db-exporter/src/converter/binary.go:59-152,
iris-mpc-common/src/helpers/inmemory_store.rs:67-73). Known answers from the
exporter's own tests (binary_test.go), and round trips at full size."""
import numpy as np
import pytest

from irismpc import store


def test_store_as_even_odd_pairs_known_answers():
    # TestStoreAsEvenOddPairs_SimpleCase and the leftCode block of
    # TestBinaryConverter_ConvertSingle
    assert store.store_as_even_odd_pairs(bytes([128, 129, 130, 131])).tolist() == [0, 2, 1, 3]
    assert store.store_as_even_odd_pairs(bytes(range(128, 136))).tolist() == [0, 2, 4, 6, 1, 3, 5, 7]
    with pytest.raises(AssertionError):  # TestStoreAsEvenOddPairs_OddLengthInput
        store.store_as_even_odd_pairs(bytes([128, 129, 130]))


def test_convert_single_known_answer():
    # TestBinaryConverter_ConvertSingle: id 42, version 55, small fields
    out = store.convert([42], [list(range(128, 136))], [[140, 141, 142, 143]],
                        [list(range(150, 158))], [[160, 161, 162, 163]], [55])
    assert list(out) == ([0, 0, 0, 42] + [0, 2, 4, 6, 1, 3, 5, 7] + [12, 14, 13, 15]
                         + [22, 24, 26, 28, 23, 25, 27, 29] + [32, 34, 33, 35] + [0, 55])


def test_record_round_trip_full_size():
    rng = np.random.default_rng(3)
    n = 3
    lc, rc = (rng.integers(0, 1 << 16, (n, 12800), dtype=np.uint16) for _ in range(2))
    lm, rm = (rng.integers(0, 1 << 16, (n, 6400), dtype=np.uint16) for _ in range(2))
    chunk = store.encode_stored_records([7, 8, 1], lc, lm, rc, rm, [1, 2, 0xFFFF])
    assert len(chunk) == 3 * store.SINGLE_ELEMENT_SIZE == 3 * 76806
    assert store.serial_ids(chunk).tolist() == [7, 8, 1]
    for i in range(n):
        rec = store.S3StoredIris.from_bytes(chunk[i * 76806:(i + 1) * 76806])
        assert rec.serial_id() == [7, 8, 1][i]
        assert rec.version_id == [1, 2, -1][i]  # i16
        got = rec.shares()
        for g, want in zip(got, (lc[i], lm[i], rc[i], rm[i])):
            assert np.array_equal(g, want)
        # "odd" = low byte limb, "even" = high byte limb, both biased by -128
        assert rec.left_code_odd[0] == (int(lc[i, 0]) & 0xFF) ^ 0x80
        assert rec.left_code_even[0] == (int(lc[i, 0]) >> 8) ^ 0x80
    with pytest.raises(ValueError):
        store.S3StoredIris.from_bytes(chunk[:1000])


def test_map_back_to_u16_is_the_limb_inverse():
    v = np.array([0, 1, 0x7F80, 0x8000, 0xFFFF, 0x1234], np.uint16)
    lo = ((v & 0xFF).astype(np.int32) - 128).astype(np.uint8)
    hi = ((v >> 8).astype(np.int32) - 128).astype(np.uint8)
    assert np.array_equal(store.map_back_to_u16(lo, hi), v)
