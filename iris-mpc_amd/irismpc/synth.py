"""Synthetic iris workloads for benchmarks and full-size tests.

Plaintext irises, planted bit flips and the three parties' GR(2^16,4) shares
come from the library's ChaCha20 generator (csrc/synth.hip); shares are
encoded on the device. A DB of n_total rows holds n_unique distinct irises
tiled (row r = iris r mod n_unique), which keeps generation cheap while the
matching work is that of the full DB.
"""
import ctypes

import numpy as np

from ._lib import call, ptr_array

CODE_WORDS = 200


def synth_iris(dataset, index):
    """Plaintext iris (dataset, index): (code, mask) as 200 u64 words each."""
    code = np.zeros(CODE_WORDS, np.uint64)
    mask = np.zeros(CODE_WORDS, np.uint64)
    call("irismpc_synth_iris", dataset, index, code.ctypes.data, mask.ctypes.data)
    return code, mask


def synth_flip(code, dataset, index, n_flips):
    """Copy of code with n_flips distinct bits flipped (draws from (dataset, index))."""
    out = np.ascontiguousarray(code, dtype=np.uint64).copy()
    call("irismpc_synth_flip", out.ctypes.data, dataset, index, n_flips)
    return out


def _shares_device(codes, masks, rand_dataset, rand_index0, stream):
    """Device shares for plaintexts held in torch u64 tensors [n][200]."""
    import torch

    n = codes.shape[0]
    cs = [torch.empty((n, 12800), dtype=torch.int16, device=codes.device) for _ in range(3)]
    ms = [torch.empty((n, 6400), dtype=torch.int16, device=codes.device) for _ in range(3)]
    call("irismpc_synth_shares_device", codes.data_ptr(), masks.data_ptr(), n, rand_dataset, rand_index0,
         ptr_array([c.data_ptr() for c in cs]), ptr_array([m.data_ptr() for m in ms]), stream)
    return cs, ms


def plaintext_device(dataset, index0, n, stream=0, device="cuda"):
    import torch

    codes = torch.empty((n, CODE_WORDS), dtype=torch.int64, device=device)
    masks = torch.empty((n, CODE_WORDS), dtype=torch.int64, device=device)
    call("irismpc_synth_plain_device", dataset, index0, n, codes.data_ptr(), masks.data_ptr(), stream)
    return codes, masks


def shares_of(codes, masks, rand_dataset, rand_index0=0, stream=0):
    """Host plaintexts [n][200] u64 -> per party (code [n][12800], mask [n][6400]) u16 on the host."""
    import torch

    c = torch.from_numpy(np.ascontiguousarray(codes, dtype=np.uint64).view(np.int64)).cuda()
    m = torch.from_numpy(np.ascontiguousarray(masks, dtype=np.uint64).view(np.int64)).cuda()
    cs, ms = _shares_device(c, m, rand_dataset, rand_index0, stream)
    torch.cuda.synchronize()
    return [(cs[p].cpu().numpy().view(np.uint16), ms[p].cpu().numpy().view(np.uint16)) for p in range(3)]


def build_db(code_dbs, mask_dbs, n_total, n_unique, dataset, rand_dataset, batch=4096, stream=0):
    """Fills each party's code/mask DB share (lists of 3 SlicedProcessedDatabase)
    with rows r = shares of plaintext iris (dataset, r mod n_unique), shared with
    randomness (rand_dataset, r mod n_unique)."""
    import torch

    assert n_unique % 16 == 0 and n_unique > 0
    n_unique = min(n_unique, -(-n_total // 16) * 16)

    def load(index0, n, iris0):
        codes, masks = plaintext_device(dataset, iris0, n, stream)
        cs, ms = _shares_device(codes, masks, rand_dataset, iris0, stream)
        for p in range(3):
            code_dbs[p].load_records_device(index0, n, cs[p].data_ptr(), stream)
            mask_dbs[p].load_records_device(index0, n, ms[p].data_ptr(), stream)
        torch.cuda.synchronize()

    first = min(n_unique, n_total)
    for i0 in range(0, first, batch):
        load(i0, min(batch, first - i0), i0)
    done = first
    while done + n_unique <= n_total:
        for p in range(3):
            code_dbs[p].replicate_records(0, n_unique, done, stream)
            mask_dbs[p].replicate_records(0, n_unique, done, stream)
        done += n_unique
    rest = n_total - done
    full = rest // 16 * 16
    if full:
        for p in range(3):
            code_dbs[p].replicate_records(0, full, done, stream)
            mask_dbs[p].replicate_records(0, full, done, stream)
        done += full
    if done < n_total:  # last rows: rows done.. hold irises done % n_unique ..
        load(done, n_total - done, done % n_unique)
    torch.cuda.synchronize()


def make_queries(batch, planted, dataset, query_dataset, flip_dataset, rand_dataset):
    """Plaintext queries: query j is a flipped copy of DB iris t for (j, t, n_flips)
    in planted, else iris (query_dataset, j). Returns (codes, masks) [batch][200]
    and per party raw shares [(code [batch][12800], mask [batch][6400])]."""
    plant = {j: (t, f) for j, t, f in planted}
    codes = np.zeros((batch, CODE_WORDS), np.uint64)
    masks = np.zeros((batch, CODE_WORDS), np.uint64)
    for j in range(batch):
        if j in plant:
            t, f = plant[j]
            c, m = synth_iris(dataset, t)
            c = synth_flip(c, flip_dataset, j, f)
        else:
            c, m = synth_iris(query_dataset, j)
        codes[j], masks[j] = c, m
    return codes, masks, shares_of(codes, masks, rand_dataset)
