"""Golden vectors of the CPU-runnable configuration, written by the oracle:
DB of 4,096 synthetic irises (12,800-bit codes, bits iid Bernoulli(0.5);
masks 1 with probability 0.85), 32 queries x 31 rotations, 8 of them copies
of DB irises with exactly 1,920 (15 %) code bits flipped, the targets drawn
from the ChaCha20 stream of the all-zero key; GR(2^16,4) shares of the three
parties; engine seeds derive_seed(([p; 8], [p - 1; 8]), nonce 0 / 1 / 5).

Kept in tests/golden/config1.npz: the planted targets, the 0.375 match bitmap
(query q, DB row d at bit 4,096 q + d), the first 8 DB columns of every
party's masked code and mask dot shares, and the first 512 words of every
party's circuit result shares (a, b).

Run from the repository root: python3 tests/golden/make_config1.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, B, FLIPS, A = 4096, 32, 1920, 16384
PLANTED = list(range(0, B, 4))   # query positions of the planted copies
COLS, WORDS = 8, 512             # dot-share columns and result words kept
CHUNK = 1 << 15
PATH = os.path.join(HERE, "config1.npz")


def planted_targets(orc):
    """Distinct DB rows from the ChaCha20 keystream of the all-zero key (nonce 0xC1)."""
    ks = orc.chacha_keystream(np.zeros(8, np.uint32), 0xC1, 0, 20, 256)
    out = []
    for w in ks:
        if int(w) % N not in out:
            out.append(int(w) % N)
        if len(out) == len(PLANTED):
            return out
    raise RuntimeError("keystream too short")


def inputs(orc):
    db_c, db_m = map(np.array, zip(*[orc.synth_iris(1, i) for i in range(N)]))
    targets = planted_targets(orc)
    q_c, q_m = [], []
    for j in range(B):
        if j in PLANTED:
            t = targets[PLANTED.index(j)]
            q_c.append(orc.synth_flip(db_c[t], 11, j, FLIPS))
            q_m.append(db_m[t])
        else:
            c, m = orc.synth_iris(21, j)
            q_c.append(c)
            q_m.append(m)
    return db_c, db_m, np.array(q_c), np.array(q_m), targets


def shares(orc, codes, masks, ds):
    cs = np.zeros((3, len(codes), 12800), np.uint16)
    ms = np.zeros((3, len(codes), 6400), np.uint16)
    for i, (c, m) in enumerate(zip(codes, masks)):
        cs[:, i], ms[:, i] = orc.synth_shares(c, m, ds, i)
    return cs, ms


def seeds(orc, p, nonce):
    return (orc.derive_seed(np.full(8, p, np.uint32), nonce),
            orc.derive_seed(np.full(8, (p + 2) % 3, np.uint32), nonce))


def expected(orc, threads=8):
    db_c, db_m, q_c, q_m, targets = inputs(orc)
    dbc_sh, dbm_sh = shares(orc, db_c, db_m, 30)
    qc_sh, qm_sh = shares(orc, q_c, q_m, 31)
    hits = np.array([orc.plain_match(q_c[j], q_m[j], db_c, db_m, A, threads) for j in range(B)])
    Qr = B * 31
    code_sh, mask_sh = [], []
    for p in range(3):
        rc = np.zeros((Qr, 12800), np.uint16)
        rm = np.zeros((Qr, 6400), np.uint16)
        for j in range(B):
            rc[j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qc_sh[p, j]))
            rm[j * 31:(j + 1) * 31] = orc.share_rotations(orc.preprocess_query(p, qm_sh[p, j]))
        sc, sm = seeds(orc, p, 0), seeds(orc, p, 1)
        code_sh.append(orc.mask_dot_share(orc.u16_dot(rc, dbc_sh[p], threads).reshape(-1), 1, sc[0], sc[1], 0, 0))
        mask_sh.append(orc.mask_dot_share(orc.u16_dot(rm, dbm_sh[p], threads).reshape(-1), 2, sm[0], sm[1], 0, 0))
    so = np.array([seeds(orc, p, 5)[0] for p in range(3)])
    sp = np.array([seeds(orc, p, 5)[1] for p in range(3)])
    ra, rb, _ = orc.threshold(code_sh, [code_sh[(p + 2) % 3] for p in range(3)], mask_sh,
                              [mask_sh[(p + 2) % 3] for p in range(3)], Qr * CHUNK // 64, A, so, sp)
    return {
        "targets": np.array(targets, np.uint32),
        "bitmap": np.packbits(hits.reshape(-1), bitorder="little").view(np.uint64),
        "code_dot_cols": np.array([c.reshape(Qr, N)[:, :COLS] for c in code_sh]),
        "mask_dot_cols": np.array([m.reshape(Qr, N)[:, :COLS] for m in mask_sh]),
        "result_a": np.array([x[:WORDS] for x in ra]),
        "result_b": np.array([x[:WORDS] for x in rb]),
    }


def load():
    with np.load(PATH) as f:
        return {k: f[k] for k in f.files}


if __name__ == "__main__":
    from oracle import orc

    orc.set_threads(min(16, os.cpu_count() or 1))
    np.savez_compressed(PATH, **expected(orc, threads=min(16, os.cpu_count() or 1)))
    print("wrote", PATH, {k: v.shape for k, v in load().items()})
