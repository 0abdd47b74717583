"""numpy/ctypes binding of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
Function-level citations of the reference live in oracle/oracle.h and src/*.c.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

CODE_LEN = 12800
MASK_LEN = 6400
ROTATIONS = 31


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, sz, u64, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    L.orc_chacha_block.argtypes = [vp, vp, i32]
    L.orc_chacha_keystream.argtypes = [vp, u64, u64, i32, vp, sz]
    L.orc_set_threads.argtypes = [i32]
    L.orc_mpc_predicate_many.argtypes = [vp, vp, sz, u32, vp]
    L.orc_plain_match.argtypes = [vp, vp, vp, vp, sz, u32, vp, i32]
    L.orc_sha256.argtypes = [vp, sz, vp]
    L.orc_hmac_sha256.argtypes = [vp, sz, vp, sz, vp]
    L.orc_hkdf_sha256.argtypes = [vp, sz, vp, sz, vp, sz, vp, sz]
    L.orc_derive_seed.argtypes = [vp, u64, vp]
    L.orc_gr_mul.argtypes = [vp, vp, vp]
    L.orc_gr_inv.argtypes = [vp, vp]
    L.orc_gr_inv.restype = i32
    L.orc_gr_A_to_monomial.argtypes = [vp, vp]
    L.orc_gr_monomial_to_B.argtypes = [vp, vp]
    L.orc_gr_lagrange_at_zero.argtypes = [vp]
    L.orc_remap_new_to_old.argtypes = [sz]
    L.orc_remap_new_to_old.restype = sz
    L.orc_remap_old_to_new.argtypes = [sz]
    L.orc_remap_old_to_new.restype = sz
    L.orc_remap_new_to_mirrored.argtypes = [sz]
    L.orc_remap_new_to_mirrored.restype = sz
    L.orc_share_mirrored.argtypes = [vp, sz, i32, vp]
    L.orc_iris_mirrored.argtypes = [vp, vp, vp, vp]
    L.orc_encode_iris_code.argtypes = [vp, vp, vp, vp]
    L.orc_encode_mask_trimmed.argtypes = [vp, vp, vp]
    L.orc_preprocess_query.argtypes = [i32, vp, sz]
    L.orc_share_rotations.argtypes = [vp, sz, vp]
    L.orc_trick_dot.argtypes = [vp, vp, sz]
    L.orc_trick_dot.restype = ctypes.c_uint16
    L.orc_iris_from_base64.argtypes = [ctypes.c_char_p, vp]
    L.orc_iris_from_base64.restype = i32
    L.orc_iris_rotate_left.argtypes = [vp, i32]
    L.orc_iris_all_rotations.argtypes = [vp, vp]
    L.orc_iris_distance.argtypes = [vp, vp, vp, vp, vp, vp]
    L.orc_mpc_predicate.argtypes = [ctypes.c_uint16, ctypes.c_uint16, u32]
    L.orc_mpc_predicate.restype = i32
    L.orc_limbs.argtypes = [vp, sz, vp, vp]
    L.orc_limb_sum.argtypes = [vp, sz]
    L.orc_limb_sum.restype = u32
    L.orc_u16_dot.argtypes = [vp, vp, sz, sz, sz, vp, i32]
    L.orc_limb_dot_corrected.argtypes = [vp, vp, sz, sz, sz, ctypes.c_uint16, vp]
    L.orc_cpu_rotation_aware_dot.argtypes = [vp, vp, vp, vp, sz, vp, i32]
    L.orc_mask_dot_share.argtypes = [vp, sz, ctypes.c_uint16, vp, vp, u64, u64]
    L.orc_threshold.argtypes = [vp, vp, vp, vp, sz, sz, u32, vp, vp, u64, u64, vp, vp, vp]
    L.orc_threshold.restype = u64
    L.orc_lift.argtypes = [vp, vp, sz, sz, vp, vp, u64, u64, vp, vp]
    L.orc_lift.restype = u64
    L.orc_open_bitmap.argtypes = [vp, vp, vp, sz, sz, sz, sz, sz, sz, vp]
    L.orc_synth_iris.argtypes = [u64, u64, vp, vp]
    L.orc_synth_flip.argtypes = [vp, u64, u64, u32, vp]
    L.orc_synth_share_rand.argtypes = [u64, u64, vp, vp]
    L.orc_synth_shares.argtypes = [vp, vp, u64, u64, vp, vp]


def set_threads(n):
    """Threads the keystream and the protocol simulator use (results do not
    depend on it; default 1)."""
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------- crypto
def chacha_block(state16, rounds):
    s = np.ascontiguousarray(state16, dtype=np.uint32)
    out = np.zeros(16, np.uint32)
    lib().orc_chacha_block(_p(s), _p(out), rounds)
    return out


def chacha_keystream(key8, nonce, counter, rounds, n_words):
    k = np.ascontiguousarray(key8, dtype=np.uint32)
    out = np.zeros(n_words, np.uint32)
    lib().orc_chacha_keystream(_p(k), nonce, counter, rounds, _p(out), n_words)
    return out


def sha256(msg: bytes) -> bytes:
    m = np.frombuffer(msg, np.uint8).copy() if msg else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().orc_sha256(_p(m), len(msg), _p(out))
    return out.tobytes()


def hkdf_sha256(salt: bytes, ikm: bytes, info: bytes, length: int) -> bytes:
    s = np.frombuffer(salt or b"\0", np.uint8).copy()
    i = np.frombuffer(ikm or b"\0", np.uint8).copy()
    f = np.frombuffer(info or b"\0", np.uint8).copy()
    out = np.zeros(length, np.uint8)
    lib().orc_hkdf_sha256(_p(s), len(salt), _p(i), len(ikm), _p(f), len(info), _p(out), length)
    return out.tobytes()


def derive_seed(seed8, nonce):
    s = np.ascontiguousarray(seed8, dtype=np.uint32)
    out = np.zeros(8, np.uint32)
    lib().orc_derive_seed(_p(s), nonce, _p(out))
    return out


# ---------------------------------------------------------------- GR(2^16,4)
def gr_mul(a, b):
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    o = np.zeros(4, np.uint16)
    lib().orc_gr_mul(_p(a), _p(b), _p(o))
    return o


def gr_inv(a):
    a = np.ascontiguousarray(a, np.uint16)
    o = np.zeros(4, np.uint16)
    rc = lib().orc_gr_inv(_p(a), _p(o))
    return o if rc == 0 else None


def gr_A_to_monomial(v):
    v = np.ascontiguousarray(v, np.uint16)
    o = np.zeros(4, np.uint16)
    lib().orc_gr_A_to_monomial(_p(v), _p(o))
    return o


def gr_monomial_to_B(v):
    v = np.ascontiguousarray(v, np.uint16)
    o = np.zeros(4, np.uint16)
    lib().orc_gr_monomial_to_B(_p(v), _p(o))
    return o


def lagrange_at_zero():
    o = np.zeros(12, np.uint16)
    lib().orc_gr_lagrange_at_zero(_p(o))
    return o.reshape(3, 4)


def remap_new_to_old(i):
    return lib().orc_remap_new_to_old(i)


def remap_old_to_new(i):
    return lib().orc_remap_old_to_new(i)


def encode_iris_code(code, mask, rand):
    code = np.ascontiguousarray(code, np.uint64)
    mask = np.ascontiguousarray(mask, np.uint64)
    rand = np.ascontiguousarray(rand, np.uint16)
    out = np.zeros((3, CODE_LEN), np.uint16)
    lib().orc_encode_iris_code(_p(code), _p(mask), _p(rand), _p(out))
    return out


def encode_mask_trimmed(mask, rand):
    mask = np.ascontiguousarray(mask, np.uint64)
    rand = np.ascontiguousarray(rand, np.uint16)
    out = np.zeros((3, MASK_LEN), np.uint16)
    lib().orc_encode_mask_trimmed(_p(mask), _p(rand), _p(out))
    return out


def preprocess_query(party_id, coefs):
    c = np.array(coefs, dtype=np.uint16, copy=True)
    lib().orc_preprocess_query(party_id, _p(c), c.size)
    return c


def share_rotations(coefs):
    c = np.ascontiguousarray(coefs, np.uint16)
    out = np.zeros((ROTATIONS, c.size), np.uint16)
    lib().orc_share_rotations(_p(c), c.size, _p(out))
    return out


def trick_dot(a, b):
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    return int(lib().orc_trick_dot(_p(a), _p(b), a.size))


# ---------------------------------------------------------------- plaintext
def iris_from_base64(s: str):
    out = np.zeros(200, np.uint64)
    rc = lib().orc_iris_from_base64(s.encode(), _p(out))
    if rc != 0:
        raise ValueError(f"bad base64 iris ({rc})")
    return out


def iris_bits(code):
    return np.unpackbits(np.ascontiguousarray(code, np.uint64).view(np.uint8), bitorder="little")


def iris_rotate_left(code, by):
    c = np.array(code, dtype=np.uint64, copy=True)
    lib().orc_iris_rotate_left(_p(c), by)
    return c


def iris_all_rotations(code):
    c = np.ascontiguousarray(code, np.uint64)
    out = np.zeros((ROTATIONS, 200), np.uint64)
    lib().orc_iris_all_rotations(_p(c), _p(out))
    return out


def iris_mirrored(code, mask):
    c = np.ascontiguousarray(code, np.uint64)
    m = np.ascontiguousarray(mask, np.uint64)
    oc, om = np.zeros(200, np.uint64), np.zeros(200, np.uint64)
    lib().orc_iris_mirrored(_p(c), _p(m), _p(oc), _p(om))
    return oc, om


def share_mirrored(coefs, flip_imaginary):
    a = np.ascontiguousarray(coefs, np.uint16)
    out = np.zeros_like(a)
    lib().orc_share_mirrored(_p(a), a.size, 1 if flip_imaginary else 0, _p(out))
    return out


def iris_distance(c1, m1, c2, m2):
    arrs = [np.ascontiguousarray(x, np.uint64) for x in (c1, m1, c2, m2)]
    h = np.zeros(1, np.uint32)
    m = np.zeros(1, np.uint32)
    lib().orc_iris_distance(*[_p(x) for x in arrs], _p(h), _p(m))
    return int(h[0]), int(m[0])


def mpc_predicate(code_dot, mask_dot, threshold_a):
    return bool(lib().orc_mpc_predicate(code_dot & 0xFFFF, mask_dot & 0xFFFF, threshold_a))


def mpc_predicate_many(code_dot, mask_dot, threshold_a):
    """mpc_predicate over arrays of u16 code / mask dots -> bool array."""
    c = np.ascontiguousarray(code_dot, np.uint16)
    m = np.ascontiguousarray(mask_dot, np.uint16)
    out = np.zeros(c.size, np.uint8)
    lib().orc_mpc_predicate_many(_p(c), _p(m), c.size, threshold_a, _p(out))
    return out.astype(bool)


def plain_match(q_code, q_mask, db_codes, db_masks, threshold_a, threads=16):
    """Plaintext brute force: bool per DB iris ([n][200] u64 codes / masks) that
    some rotation of the query matches at the integer MPC predicate."""
    qc = np.ascontiguousarray(q_code, np.uint64)
    qm = np.ascontiguousarray(q_mask, np.uint64)
    dc = np.ascontiguousarray(db_codes, np.uint64).reshape(-1, 200)
    dm = np.ascontiguousarray(db_masks, np.uint64).reshape(-1, 200)
    out = np.zeros(dc.shape[0], np.uint8)
    lib().orc_plain_match(_p(qc), _p(qm), _p(dc), _p(dm), dc.shape[0], threshold_a, _p(out), threads)
    return out.astype(bool)


# ---------------------------------------------------------------- dots
def limbs(v):
    v = np.ascontiguousarray(v, np.uint16)
    l0 = np.zeros(v.shape, np.int8)
    l1 = np.zeros(v.shape, np.int8)
    lib().orc_limbs(_p(v), v.size, _p(l0), _p(l1))
    return l0, l1


def limb_sum(l):
    l = np.ascontiguousarray(l, np.int8)
    return int(lib().orc_limb_sum(_p(l), l.size))


def u16_dot(q, db, threads=8):
    q = np.ascontiguousarray(q, np.uint16)
    db = np.ascontiguousarray(db, np.uint16)
    Qr, K = q.shape
    N = db.shape[0]
    out = np.zeros((Qr, N), np.uint16)
    lib().orc_u16_dot(_p(q), _p(db), Qr, N, K, _p(out), threads)
    return out


def limb_dot_corrected(q, db, multiplier=1):
    q = np.ascontiguousarray(q, np.uint16)
    db = np.ascontiguousarray(db, np.uint16)
    Qr, K = q.shape
    N = db.shape[0]
    out = np.zeros((Qr, N), np.uint16)
    lib().orc_limb_dot_corrected(_p(q), _p(db), Qr, N, K, multiplier, _p(out))
    return out


def cpu_rotation_aware_dot(q_code, q_mask, db_code, db_mask, threads=8):
    q_code = np.ascontiguousarray(q_code, np.uint16)
    q_mask = np.ascontiguousarray(q_mask, np.uint16)
    db_code = np.ascontiguousarray(db_code, np.uint16)
    db_mask = np.ascontiguousarray(db_mask, np.uint16)
    n = db_code.shape[0]
    out = np.zeros((n, ROTATIONS, 2), np.uint16)
    lib().orc_cpu_rotation_aware_dot(_p(q_code), _p(q_mask), _p(db_code), _p(db_mask), n, _p(out), threads)
    return out


def mask_dot_share(dot, multiplier, seed_own, seed_prev, nonce, counter):
    d = np.array(dot, dtype=np.uint16, copy=True)
    so = np.ascontiguousarray(seed_own, np.uint32)
    sp = np.ascontiguousarray(seed_prev, np.uint32)
    lib().orc_mask_dot_share(_p(d), d.size, multiplier, _p(so), _p(sp), nonce, counter)
    return d


# ---------------------------------------------------------------- protocol
def _ptr3(arrs):
    return (ctypes.c_void_p * 3)(*[a.ctypes.data for a in arrs])


def threshold(code_a, code_b, mask_a, mask_b, alloc_words, threshold_a, seeds_own, seeds_prev,
              nonce=0, counter=0, want_lifted=False):
    """code_a[p] etc: u16 arrays of length L for party p. Returns (res_a, res_b, ctr_end[, lifted])."""
    ca = [np.ascontiguousarray(x, np.uint16) for x in code_a]
    cb = [np.ascontiguousarray(x, np.uint16) for x in code_b]
    ma = [np.ascontiguousarray(x, np.uint16) for x in mask_a]
    mb = [np.ascontiguousarray(x, np.uint16) for x in mask_b]
    L = ca[0].size
    assert L % 64 == 0
    W = L // 64
    ra = [np.zeros(W, np.uint64) for _ in range(3)]
    rb = [np.zeros(W, np.uint64) for _ in range(3)]
    so = np.ascontiguousarray(seeds_own, np.uint32).reshape(3, 8)
    sp = np.ascontiguousarray(seeds_prev, np.uint32).reshape(3, 8)
    lifted = [np.zeros(L, np.uint32) for _ in range(3)] if want_lifted else None
    ctr = lib().orc_threshold(_ptr3(ca), _ptr3(cb), _ptr3(ma), _ptr3(mb), L, alloc_words, threshold_a,
                              _p(so), _p(sp), nonce, counter, _ptr3(ra), _ptr3(rb),
                              _ptr3(lifted) if want_lifted else None)
    if want_lifted:
        return ra, rb, ctr, lifted
    return ra, rb, ctr


def lift(mask_a, mask_b, alloc_words, seeds_own, seeds_prev, nonce=0, counter=0):
    ma = [np.ascontiguousarray(x, np.uint16) for x in mask_a]
    mb = [np.ascontiguousarray(x, np.uint16) for x in mask_b]
    L = ma[0].size
    oa = [np.zeros(L, np.uint32) for _ in range(3)]
    ob = [np.zeros(L, np.uint32) for _ in range(3)]
    so = np.ascontiguousarray(seeds_own, np.uint32).reshape(3, 8)
    sp = np.ascontiguousarray(seeds_prev, np.uint32).reshape(3, 8)
    ctr = lib().orc_lift(_ptr3(ma), _ptr3(mb), L, alloc_words, _p(so), _p(sp), nonce, counter,
                         _ptr3(oa), _ptr3(ob))
    return oa, ob, ctr


def open_bitmap(r1, r2, r3, chunk_len, query_len, offset, real_chunk_len, total_db_len, bitmap):
    r1, r2, r3 = [np.ascontiguousarray(x, np.uint64) for x in (r1, r2, r3)]
    lib().orc_open_bitmap(_p(r1), _p(r2), _p(r3), r1.size, chunk_len, query_len, offset,
                          real_chunk_len, total_db_len, _p(bitmap))
    return bitmap


# ---------------------------------------------------------------- synthetic data
def synth_iris(dataset, index):
    c = np.zeros(200, np.uint64)
    m = np.zeros(200, np.uint64)
    lib().orc_synth_iris(dataset, index, _p(c), _p(m))
    return c, m


def synth_flip(code, dataset, index, n_flips):
    c = np.ascontiguousarray(code, np.uint64)
    out = np.zeros(200, np.uint64)
    lib().orc_synth_flip(_p(c), dataset, index, n_flips, _p(out))
    return out


def synth_shares(code, mask, dataset, index):
    c = np.ascontiguousarray(code, np.uint64)
    m = np.ascontiguousarray(mask, np.uint64)
    cs = np.zeros((3, CODE_LEN), np.uint16)
    ms = np.zeros((3, MASK_LEN), np.uint16)
    lib().orc_synth_shares(_p(c), _p(m), dataset, index, _p(cs), _p(ms))
    return cs, ms
