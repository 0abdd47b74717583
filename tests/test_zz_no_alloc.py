"""irismpc_partial_results_noalloc (the candidate list without device
allocation) against irismpc_comparator_get_partial_results, through the C
ABI: two calls in a row on one GPU, the second with a smaller DB, bitmap and
extra-id list, so the second call also runs over scratch contents the first
left behind. Both must give the same ascending id list, bit for bit, and the
list must be the union of the set bitmap bits and the in-range extra ids."""
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


def _case(irismpc, torch, rng, db_size, batch, row_bits, n_set, extra):
    rb = row_bits or db_size
    words = -(-(rb * batch) // 64) + 2
    host = np.zeros(words, np.uint64)
    want = set()
    for q in range(batch):
        for d in rng.choice(db_size, size=n_set, replace=False):
            bit = q * rb + int(d)
            host[bit // 64] |= np.uint64(1) << np.uint64(bit % 64)
            want.add(int(d))
    want |= {int(d) for d in extra if d < db_size}
    bitmap = torch.from_numpy(host.view(np.int64)).cuda()
    cmp = irismpc.DistanceComparator(batch * 31, db_size)
    ref = cmp.get_partial_results(bitmap.data_ptr(), db_size, batch, extra_ids=extra, row_bits=row_bits)
    ext = torch.from_numpy(np.asarray(extra, np.uint32).view(np.int32)).cuda()
    ids = torch.full((db_size + 1,), -1, dtype=torch.int32, device="cuda")
    count = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    stream = irismpc.stream_create()
    torch.cuda.synchronize()
    irismpc.partial_results_noalloc(bitmap.data_ptr(), db_size, batch, ids.data_ptr(), count.data_ptr(),
                                    ext.data_ptr() if len(extra) else None, len(extra), row_bits, stream)
    irismpc.stream_synchronize(stream)
    n = int(count.cpu().numpy()[0])
    got = ids.cpu().numpy().view(np.uint32)
    assert np.array_equal(got[:n], ref), f"db {db_size}: no-alloc list != get_partial_results list"
    assert got[:n].tolist() == sorted(want)
    assert np.all(got[n:] == 0xFFFFFFFF), "written past the list's end"
    return n


def test_noalloc_candidates_equal_get_partial_results(irismpc):
    import torch

    rng = np.random.default_rng(27)
    n1 = _case(irismpc, torch, rng, 1_000_003, 64, 1_048_576, 40,
               [5, 5, 999_999, 1_000_002, 1_000_003, 2 ** 31, 17])
    n2 = _case(irismpc, torch, rng, 70_000, 5, 0, 10, [69_999, 3])
    assert n1 > n2 > 0
