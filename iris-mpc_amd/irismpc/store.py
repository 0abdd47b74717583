"""DB-chunk record format of the reference's store and its exporter.

A DB chunk (an S3 object) is a run of fixed-size records
(iris-mpc-store/src/s3_importer.rs:13-94), written by the DB exporter
(iris-mpc-db-exporter/src/converter/binary.go:59-152):
  big-endian u32 serial id;
  left code, left mask, right code, right mask: each share (u16 little-endian
  bytes) split into its low bytes ("odd") and then its high bytes ("even"),
  every byte biased by -128 (mod 256), which makes them the i8 limbs the
  share GEMM uses;
  big-endian u16 version id.
76,806 bytes per record. load_stored_records puts party p's records
straight into HBM (one host-to-device copy of the records; the limbs are
de-interleaved into the fragment layout on the device).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import call

IRIS_CODE_LENGTH, MASK_CODE_LENGTH = 12800, 6400
SINGLE_ELEMENT_SIZE = 4 + 2 * (2 * IRIS_CODE_LENGTH) + 2 * (2 * MASK_CODE_LENGTH) + 2  # 76,806
FIELDS = (("left_code", IRIS_CODE_LENGTH), ("left_mask", MASK_CODE_LENGTH),
          ("right_code", IRIS_CODE_LENGTH), ("right_mask", MASK_CODE_LENGTH))


def store_as_even_odd_pairs(data):
    """storeAsEvenOddPairs (binary.go:120-152): bytes at even positions
    first, then those at odd positions, each minus 128 (mod 256)."""
    b = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data.view(np.uint8)
    assert b.size % 2 == 0, "input must have an even number of bytes"
    return np.concatenate([b[0::2] - np.uint8(128), b[1::2] - np.uint8(128)])


def map_back_to_u16(odd, even):
    """inmemory_store.rs:67-73: u16 from its biased low ("odd") and high
    ("even") bytes."""
    lo = (np.asarray(odd, np.uint8) + np.uint8(128)).astype(np.uint16)
    hi = (np.asarray(even, np.uint8) + np.uint8(128)).astype(np.uint16)
    return lo | (hi << 8)


def convert(ids, left_code, left_mask, right_code, right_mask, version_ids=None):
    """BinaryConverter.Convert (binary.go:59-87) on the raw stored bytes of
    each field ([n][bytes] u8, any even size): one record per id."""
    n = len(ids)
    fields = [np.asarray(f, np.uint8).reshape(n, -1) for f in (left_code, left_mask, right_code, right_mask)]
    size = 4 + sum(f.shape[1] for f in fields) + 2
    version_ids = np.zeros(n, np.uint16) if version_ids is None else np.asarray(version_ids)
    out = np.zeros((n, size), np.uint8)
    out[:, 0:4] = np.asarray(ids, ">u4").view(np.uint8).reshape(n, 4)
    off = 4
    for f in fields:
        assert f.shape[1] % 2 == 0, "input must have an even number of bytes"
        h = f.shape[1] // 2
        out[:, off:off + h] = f[:, 0::2] - np.uint8(128)
        out[:, off + h:off + 2 * h] = f[:, 1::2] - np.uint8(128)
        off += f.shape[1]
    out[:, off:off + 2] = np.asarray(version_ids, ">u2").view(np.uint8).reshape(n, 2)
    return out.reshape(-1).tobytes()


def encode_stored_records(ids, left_code, left_mask, right_code, right_mask, version_ids=None):
    """One party's u16 share arrays ([n][12800] codes, [n][6400] masks) -> the
    chunk bytes the exporter writes (shares stored as little-endian bytes)."""
    n = len(ids)
    as_bytes = lambda a, K: np.ascontiguousarray(a, "<u2").reshape(n, K).view(np.uint8)  # noqa: E731
    return convert(ids, as_bytes(left_code, IRIS_CODE_LENGTH), as_bytes(left_mask, MASK_CODE_LENGTH),
                   as_bytes(right_code, IRIS_CODE_LENGTH), as_bytes(right_mask, MASK_CODE_LENGTH),
                   version_ids)


@dataclass
class S3StoredIris:
    """S3StoredIris::from_bytes (s3_importer.rs:34-94) of one record."""
    id: int
    left_code_odd: bytes
    left_code_even: bytes
    left_mask_odd: bytes
    left_mask_even: bytes
    right_code_odd: bytes
    right_code_even: bytes
    right_mask_odd: bytes
    right_mask_even: bytes
    version_id: int

    @classmethod
    def from_bytes(cls, b):
        if len(b) < SINGLE_ELEMENT_SIZE:
            raise ValueError("Exceeded total bytes while extracting slice")
        cur = 4
        parts = []
        for _, K in FIELDS:
            parts += [bytes(b[cur:cur + K]), bytes(b[cur + K:cur + 2 * K])]
            cur += 2 * K
        version = int.from_bytes(b[cur:cur + 2], "big")
        version = version - 65536 if version >= 32768 else version  # as i16
        return cls(int.from_bytes(b[0:4], "big"), *parts, version)

    def serial_id(self):
        return self.id

    def shares(self):
        """(left code, left mask, right code, right mask) as u16 arrays."""
        return tuple(map_back_to_u16(np.frombuffer(o, np.uint8), np.frombuffer(e, np.uint8))
                     for o, e in ((self.left_code_odd, self.left_code_even),
                                  (self.left_mask_odd, self.left_mask_even),
                                  (self.right_code_odd, self.right_code_even),
                                  (self.right_mask_odd, self.right_mask_even)))


def serial_ids(chunk):
    """Serial ids of the records of a chunk."""
    a = np.frombuffer(chunk, np.uint8)
    assert a.size % SINGLE_ELEMENT_SIZE == 0, "chunk is not a whole number of records"
    return a.reshape(-1, SINGLE_ELEMENT_SIZE)[:, :4].copy().view(">u4").reshape(-1).astype(np.uint32)


def load_stored_records(codes, masks, left_code, left_mask, right_code, right_mask, chunk,
                        max_serial_id=0, stream=0):
    """One party's chunk bytes -> its four HBM DB slices (record with serial id
    s at row s - 1; ids above max_serial_id skipped, 0 = no cap). Returns
    (records loaded, largest loaded row + 1)."""
    buf = np.frombuffer(chunk, np.uint8)
    assert buf.size % SINGLE_ELEMENT_SIZE == 0, "chunk is not a whole number of records"
    n_loaded, row_end = ctypes.c_size_t(), ctypes.c_size_t()
    call("irismpc_db_load_stored_records", codes.h, masks.h, left_code.h, left_mask.h, right_code.h,
         right_mask.h, buf.ctypes.data if buf.size else None, buf.size // SINGLE_ELEMENT_SIZE,
         max_serial_id, ctypes.byref(n_loaded), ctypes.byref(row_end), stream)
    return n_loaded.value, row_end.value
