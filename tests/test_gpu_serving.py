"""Serving form: continuous batching over the DB scan, with reauth, insertion
and deletion requests in the mix.

Checked on a 4-chunk DB of synthetic irises (4,096 distinct irises tiled):
- every uniqueness result in the mix equals the result of the same request
  alone (the plain per-request form) and the answer known by construction
  (planted copies of DB iris t match exactly rows r with r mod 4096 = t);
- reauth against the right / a wrong row; an inserted iris is found by a
  later query; a deleted row no longer matches;
- requests overlap in the active set; admission by free HBM lets one request
  run at a time when the reserve leaves no room;
- step sizes chosen by the autotuner leave every result unchanged;
- a server starting from an empty DB (as the reference's does) answers
  uniqueness requests with no matches, then finds an inserted iris.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, U = 4 * 32768, 4096


@pytest.fixture(scope="module")
def served():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import irismpc
    from irismpc import serving, synth

    irismpc.lib()
    eng = serving.ServingEngines(N, N + 2 * 32768, max_active=8, seed_base=5)
    eng.build(U, 71, 72)
    return eng, serving, synth


def _query(synth, ds_iris, t, flip_idx, flips, rand_idx):
    c, m = synth.synth_iris(ds_iris, t)
    if flips:
        c = synth.synth_flip(c, 75, flip_idx, flips)
    sh = synth.shares_of(c[None], m[None], 76, rand_idx)
    return [(sh[p][0][0], sh[p][1][0]) for p in range(3)], c, m


def test_serving_mix_matches_alone(served):
    eng, serving, synth = served
    R = serving.Request
    rng = np.random.default_rng(3)
    reqs, expect = [], {}
    targets = rng.choice(U, size=6, replace=False)
    deleted = int(targets[2]) + 2 * U  # one copy of iris targets[2], deleted at t = 0
    # uniqueness: 6 planted (10-30 % flips), 4 fresh irises
    for k in range(10):
        if k < 6:
            sh, _, _ = _query(synth, 71, int(targets[k]), k, int(rng.integers(1280, 3841)), k)
            expect[k] = [r for r in range(N) if r % U == targets[k] and r != deleted]
        else:
            sh, _, _ = _query(synth, 73, k, 0, 0, k)
            expect[k] = []
        reqs.append(R("uniqueness", 0.002 * k, shares=sh, rid=k))
    # reauth: right row, wrong row
    sh, _, _ = _query(synth, 71, int(targets[0]), 50, 1500, 50)
    reqs.append(R("reauth", 0.001, shares=sh, target=int(targets[0]) + U, rid=20))
    reqs.append(R("reauth", 0.003, shares=sh, target=int(targets[1]), rid=21))
    # insertion of a fresh iris, deletion of one copy of targets[2]
    new_c, new_m = synth.synth_iris(74, 1)
    ins = synth.shares_of(new_c[None], new_m[None], 77, 1)
    reqs.append(R("insert", 0.0, shares=[(ins[p][0][0], ins[p][1][0]) for p in range(3)], rid=30))
    reqs.append(R("delete", 0.0, target=deleted, rid=31))
    # after both applied: a query for the inserted iris, one for the deleted row's iris
    sh_new, _, _ = _query(synth, 74, 1, 60, 2000, 60)
    sh_del, _, _ = _query(synth, 71, int(targets[2]), 61, 2000, 61)
    reqs.append(R("uniqueness", 0.6, shares=sh_new, rid=40))
    reqs.append(R("uniqueness", 0.6, shares=sh_del, rid=41))

    srv = serving.Server(eng, reserve_bytes=0)
    out = {r.rid: r for r in srv.run(reqs)}
    assert srv.max_concurrent >= 4, "requests did not overlap in the active set"
    for k in range(10):
        assert out[k].result == expect[k], f"uniqueness {k} in the mix"
    assert out[20].result is True and out[21].result is False
    assert out[30].result == N and eng.n_rows == N + 1
    assert out[40].result == [N]
    assert out[41].result == expect[2]
    # the same requests alone (plain per-request form)
    for k in (0, 3, 7):
        assert eng.run_alone(reqs[k].shares) == out[k].result, f"uniqueness {k} alone"
    assert eng.run_alone(sh_new) == [N]


def test_admission_by_free_hbm(served):
    import torch

    eng, serving, synth = served
    R = serving.Request
    reqs = []
    for k in range(4):
        sh, _, _ = _query(synth, 73, 100 + k, 0, 0, 100 + k)
        reqs.append(R("uniqueness", 0.0, shares=sh, rid=k))
    free, _ = torch.cuda.mem_get_info()
    srv = serving.Server(eng, reserve_bytes=free + (1 << 30))  # no room beyond the first request
    out = srv.run(reqs)
    assert all(r.result == [] for r in out)
    assert srv.max_concurrent == 1 and srv.admission_waits > 0


def test_tuned_step_sizes_keep_results(served, tmp_path):
    eng, serving, synth = served
    from irismpc.batching import Autotuner

    sh, _, _ = _query(synth, 71, 5, 90, 2000, 90)
    plain = eng.run_alone(sh)
    assert len(plain) >= N // U - 1 and all(r % U == 5 for r in plain)  # one copy may be deleted
    table = eng.tune_steps(Autotuner(str(tmp_path / "tune.json")))
    try:
        assert set(table) == {1, 2, 4, 8}  # buckets up to max_active = 8
        for b, k in table.items():
            assert k >= 1 and k * b <= eng.max_active and k <= eng.n_chunks(eng.n_rows)
        assert eng.run_alone(sh) == plain
        # cached: a second tuner on the same file returns the same table without timing
        assert Autotuner(str(tmp_path / "tune.json")).step_chunks(eng) == table
    finally:
        eng.step_table = None


def test_empty_db_then_insert(served):
    _, serving, synth = served
    import irismpc

    eng = serving.ServingEngines(0, 2 * irismpc.DB_CHUNK_SIZE, max_active=2, seed_base=9)
    sh, _, _ = _query(synth, 78, 3, 0, 0, 3)
    assert eng.run_alone(sh) == []
    R = serving.Request
    out = serving.Server(eng, reserve_bytes=0).run([R("uniqueness", 0.0, shares=sh, rid=0),
                                                    R("insert", 0.0, shares=sh, rid=1)])
    res = {r.rid: r for r in out}
    assert res[1].result == 0 and eng.n_rows == 1
    assert res[0].result in ([], [0])  # checked before or after the insertion, as it was admitted
    near, _, _ = _query(synth, 78, 3, 5, 1500, 4)  # the inserted iris with 1,500 code bits flipped
    assert eng.run_alone(near) == [0]
