"""Serving form: a request loop with continuous batching over the DB scan,
and the reference's per-batch form on the same engines.

Both loops are native (the C ABI's irismpc_serving_*: a loop thread and a
reauth worker over the three parties' engines); this module keeps the
host-side names used by the tests and the benchmark.

Requests arrive over time:
- uniqueness: one query x 31 rotations against the whole DB;
- reauth: one query x 31 rotations against one target row;
- insertion: a new iris appended to the DB;
- deletion: a DB row overwritten with the shares of the default iris, as
  get_dummy_shares_for_deletion (iris-mpc-common/src/iris_db/mod.rs:10-23)
  does; code all zeros and mask all ones never reaches the threshold.

The three parties' DB shares and protocol state stay on the device. The scan
of actor.rs:2246-2589 becomes a ring of chunk steps (32,768 rows each):
- Uniqueness requests join the active query set between chunk steps, as slots
  free up and while free HBM stays above a reserve.
- A request leaves when it has covered every chunk of the DB, so requests
  that arrive during a scan start at the next step instead of waiting for
  the scan to finish. A chunk written (insertion, deletion) while requests
  are active is scanned again for all of them at once, so every answer is
  that of the DB at the request's last step: after exactly
  Request.db_version of the loop's writes, in the order it applied them.
- Every step runs one share GEMM per matrix and party for all active queries,
  then the reshare, the threshold circuits and the open into each request's
  bitmap row (slot map).
- Requests in flight together are compared with each other, as the
  reference compares a batch with itself (compare_query_against_self,
  mergeBatchResults): the newly admitted requests against the active set's
  irises, on a third engine set beside the scan; Request.batch_matches lists
  the earlier-admitted requests (rids) whose iris a request matches.
- Reauth runs on its own small engine set and streams, concurrently with the
  scan.
- Insertions and deletions are written between steps, ordered on the scan
  streams.

The per-batch form (PerBatchServer) runs the reference's batch after batch
(actor.rs:584-644): every batch padded to max_batch x 31 query rows
(actor.rs:895-901), one full scan per batch, each batch compared with itself.

The opened bits do not depend on the mix (the masks do), so each result
equals the result of the same request run alone on the same DB version
(ServingEngines.run_alone, the plain per-request form).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import engine as E
from . import synth
from ._lib import call

ROT = E.ROTATIONS
CHUNK = E.DB_CHUNK_SIZE
KINDS = {"uniqueness": 0, "reauth": 1, "insert": 2, "delete": 3}
CONTINUOUS, PER_BATCH = 0, 1
PENDING, DONE, FAILED = 0, 1, -1


@dataclass
class Request:
    kind: str                 # "uniqueness" | "reauth" | "insert" | "delete"
    arrival: float            # seconds from the start of the run
    shares: list = None       # per party (code [12800] u16, mask [6400] u16): query or new iris
    target: int = -1          # reauth target row / deletion row
    rid: int = 0
    result: object = None     # uniqueness: sorted matching rows; reauth: bool; insert: row; delete: row
    done: float = -1.0
    error: str = None         # set when serving the request failed
    db_rows: int = 0          # DB rows the request was checked against
    db_version: int = 0       # DB writes the answer reflects (a write: its own position, from 1)
    # uniqueness: the requests admitted before it and scanning with it (its
    # batch, in the per-batch form) whose iris it matches (native ids from
    # fetch; the request loops map them to the Request objects' rids)
    batch_matches: list = None


def _pack_shares(shares):
    """Per-party (code, mask) -> contiguous [3][12800] and [3][6400] u16."""
    code = np.ascontiguousarray(np.stack([np.asarray(shares[p][0], np.uint16) for p in range(3)]))
    mask = np.ascontiguousarray(np.stack([np.asarray(shares[p][1], np.uint16) for p in range(3)]))
    assert code.shape == (3, E.IRIS_CODE_LENGTH) and mask.shape == (3, E.MASK_CODE_LENGTH)
    return code, mask


class ServingEngines:
    """The native request loop's state: the three parties' DB shares, the scan
    engine set (max_active queries) and the reauth engine set (one query).
    dbs = (code_dbs, mask_dbs), per party, serves DB slices that already hold
    n_rows rows and room for `capacity` (written by insertions and deletions);
    otherwise the slices are allocated here."""

    def __init__(self, n_rows, capacity, max_active=64, seed_base=0, threshold_a=None, dbs=None):
        self.capacity, self.max_active = capacity, max_active
        self.A = E.THRESHOLD_ANON_STATS_A if threshold_a is None else threshold_a
        if dbs is None:
            zero = np.zeros(8, np.uint32)
            self._loaders = [(E.ShareDB(p, CHUNK, ROT, E.IRIS_CODE_LENGTH, (zero, zero)),
                              E.ShareDB(p, CHUNK, ROT, E.MASK_CODE_LENGTH, (zero, zero))) for p in range(3)]
            self.code_dbs = [self._loaders[p][0].alloc_db(capacity) for p in range(3)]
            self.mask_dbs = [self._loaders[p][1].alloc_db(capacity) for p in range(3)]
        else:
            self.code_dbs, self.mask_dbs = list(dbs[0]), list(dbs[1])
        seeds = np.ascontiguousarray(np.stack([np.full(8, seed_base + p, np.uint32) for p in range(3)]))
        h = ctypes.c_void_p()
        call("irismpc_serving_create", 0, seeds.ctypes.data, E.ptr_array([d.h for d in self.code_dbs]),
             E.ptr_array([d.h for d in self.mask_dbs]), n_rows, capacity, max_active, self.A, ctypes.byref(h))
        self.h = h.value
        self._table = None

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h and call is not None:  # module globals are gone at interpreter exit
            call("irismpc_serving_destroy", h)

    @property
    def n_rows(self):
        n = ctypes.c_size_t()
        call("irismpc_serving_n_rows", self.h, ctypes.byref(n))
        return n.value

    def build(self, n_unique, dataset, rand_dataset):
        synth.build_db(self.code_dbs, self.mask_dbs, self.n_rows, n_unique, dataset, rand_dataset)

    @staticmethod
    def n_chunks(rows):
        return max(1, -(-rows // CHUNK))

    def chunks_per_step(self, batch, chunk_index, n_chunks):
        k = ctypes.c_size_t()
        call("irismpc_serving_chunks_per_step", self.h, batch, chunk_index, n_chunks, ctypes.byref(k))
        return k.value

    @property
    def step_table(self):
        """{active-query bucket: chunks per step} in use (None: the fixed rule)."""
        return None if self._table is None else dict(self._table)

    @step_table.setter
    def step_table(self, table):
        items = sorted((table or {}).items())
        b = (ctypes.c_uint32 * max(1, len(items)))(*[int(x) for x, _ in items])
        k = (ctypes.c_uint32 * max(1, len(items)))(*[int(y) for _, y in items])
        call("irismpc_serving_set_step_table", self.h, b, k, len(items))
        self._table = dict(items) if items else None

    def set_admission(self, active_cap=None, queue_threshold=0):
        """Serving-form admission: up to active_cap active requests, all waiting
        ones while more than queue_threshold wait (None: every free slot)."""
        call("irismpc_serving_set_admission", self.h, active_cap or self.max_active, queue_threshold)

    def time_step(self, batch, n_chunks, reps=2):
        """Seconds of one scan step of `batch` default-iris queries over n_chunks chunks."""
        t = ctypes.c_double()
        call("irismpc_serving_time_step", self.h, batch, n_chunks, reps, ctypes.byref(t))
        return t.value

    def tune_steps(self, tuner=None, **kw):
        """Step sizes per active-query count from the autotuner (batching.Autotuner)."""
        from .batching import Autotuner

        self.step_table = (tuner or Autotuner()).step_chunks(self, **kw)
        return self.step_table

    def run_alone(self, shares):
        """The plain per-request form: a full scan of one uniqueness query."""
        code, mask = _pack_shares(shares)
        cap = max(1, self.n_rows)  # every row at most once
        out, n = np.empty(cap, np.uint32), ctypes.c_size_t()
        call("irismpc_serving_run_alone", self.h, code.ctypes.data, mask.ctypes.data, out.ctypes.data, cap,
             ctypes.byref(n))
        return out[:n.value].tolist()

    # ------------------------------------------------ the native request loop
    def now(self):
        t = ctypes.c_double()
        call("irismpc_serving_now", self.h, ctypes.byref(t))
        return t.value

    def start(self, policy, max_batch=0, reserve_bytes=0):
        call("irismpc_serving_start", self.h, policy, max_batch, reserve_bytes)

    def finish(self, timeout_s=0.0):
        call("irismpc_serving_finish", self.h, float(timeout_s))

    def submit(self, r, arrival):
        i = ctypes.c_uint64()
        if r.kind == "delete":
            code = mask = None
            cp = mp = None
        else:
            code, mask = _pack_shares(r.shares)
            cp, mp = code.ctypes.data, mask.ctypes.data
        call("irismpc_serving_submit", self.h, KINDS[r.kind], cp, mp, int(r.target), float(arrival),
             ctypes.byref(i))
        return i.value

    def fetch(self, i, r, t0):
        """Copies request i's outcome into r (times relative to t0)."""
        state, value = ctypes.c_int(), ctypes.c_int64()
        times, rows = (ctypes.c_double * 2)(), ctypes.c_size_t()
        err = ctypes.create_string_buffer(512)
        n = ctypes.c_size_t()
        call("irismpc_serving_result", self.h, i, ctypes.byref(state), ctypes.byref(value), None, 0,
             ctypes.byref(n), times, ctypes.byref(rows), err, len(err))
        r.done, r.db_rows = times[1] - t0, rows.value
        if state.value == FAILED:
            r.error = err.value.decode(errors="replace")
            return
        assert state.value == DONE, f"request {r.kind} rid {r.rid} unfinished"
        version = ctypes.c_uint64()
        call("irismpc_serving_result_version", self.h, i, ctypes.byref(version))
        r.db_version = version.value
        if r.kind == "uniqueness":
            out = np.zeros(max(1, n.value), np.uint32)
            if n.value:
                call("irismpc_serving_result", self.h, i, None, None, out.ctypes.data, n.value, ctypes.byref(n),
                     None, None, None, 0)
            r.result = out[:n.value].tolist()
            k = ctypes.c_size_t()
            call("irismpc_serving_result_batch_matches", self.h, i, None, 0, ctypes.byref(k))
            ids = (ctypes.c_uint64 * max(1, k.value))()
            if k.value:
                call("irismpc_serving_result_batch_matches", self.h, i, ids, k.value, ctypes.byref(k))
            r.batch_matches = list(ids[:k.value])
        elif r.kind == "reauth":
            r.result = bool(value.value)
        else:
            r.result = int(value.value)

    def stats(self):
        out = (ctypes.c_uint64 * 5)()
        call("irismpc_serving_stats", self.h, out)
        t = (ctypes.c_double * 5)()
        call("irismpc_serving_loop_times", self.h, t)
        return {"steps": out[0], "max_concurrent": out[1], "admission_waits": out[2], "batches": out[3],
                "refills": out[4], "loop_seconds": {"refill": t[0], "wait_queued_steps": t[1],
                                                    "enqueue_steps": t[2], "free_hbm_checks": t[3],
                                                    "decode": t[4]}}


class _Loop:
    """Runs `requests` (arrival: seconds from the start) through the native
    loop in real time: the whole schedule is submitted up front, the loop
    takes each request at its arrival time, and the call returns when all
    have finished. The start lies a little after the call (time to hand the
    schedule over), so no request arrives before it is submitted."""

    def __init__(self, eng, policy, max_batch=0, reserve_bytes=0, timeout_s=120.0):
        self.eng, self.policy, self.max_batch, self.reserve = eng, policy, max_batch, reserve_bytes
        self.timeout_s = timeout_s
        self.failed = []

    def run(self, requests):
        eng = self.eng
        reqs = sorted(requests, key=lambda r: r.arrival)
        eng.start(self.policy, self.max_batch, self.reserve)
        t0 = eng.now() + 0.05 + 2e-4 * len(reqs)
        ids = []
        try:
            for r in reqs:
                ids.append(eng.submit(r, t0 + r.arrival))
            late = eng.now() - t0
            if late > 0:
                raise RuntimeError(f"the schedule was handed over {late * 1e3:.1f} ms after its start")
        finally:
            eng.finish(self.timeout_s + max(0.0, reqs[-1].arrival if reqs else 0.0))
        rid_of = {i: r.rid for r, i in zip(reqs, ids)}
        for r, i in zip(reqs, ids):
            eng.fetch(i, r, t0)
            call("irismpc_serving_release", eng.h, i)
            if r.batch_matches is not None:
                r.batch_matches = [rid_of[j] for j in r.batch_matches]
            if r.error:
                self.failed.append(r)
        st = self.stats = eng.stats()
        self.steps, self.max_concurrent = st["steps"], st["max_concurrent"]
        self.admission_waits, self.batches = st["admission_waits"], st["batches"]
        if self.failed:
            r = self.failed[0]
            raise RuntimeError(f"{len(self.failed)} request(s) failed; first: {r.kind} rid {r.rid}: {r.error}")
        return reqs


class Server(_Loop):
    """The serving form's request loop over ServingEngines. Admission: a
    uniqueness request takes a free slot while free HBM stays above
    `reserve_bytes` plus its device footprint."""

    def __init__(self, eng, reserve_bytes=4 << 30):
        super().__init__(eng, CONTINUOUS, 0, reserve_bytes)


class PerBatchServer(_Loop):
    """The plain per-batch form of the reference's server (batch after batch,
    actor.rs:584-644): the uniqueness and reauth requests that arrive while a
    batch runs form the next batch (at most max_batch of them); every batch is
    one scan of the whole DB whose query set is padded to max_batch x 31 rows,
    as the reference sizes its queries and GEMMs by max_batch_size whatever
    the batch holds (actor.rs:895-901, share_db.rs:800-845 with query_length
    = max_batch_size x 31). A reauth is answered by its row of the batch's scan
    at its target; insertions and deletions that arrived before a batch starts
    are applied first. Same engines, DB and results as the serving form."""

    def __init__(self, eng, max_batch):
        assert 1 <= max_batch <= eng.max_active
        super().__init__(eng, PER_BATCH, max_batch, 0)


def poisson_mix(rate, n, rng, make_uniqueness, make_reauth, make_insert, make_delete,
                mix=(0.85, 0.10, 0.05)):
    """n requests with exponential inter-arrival times at `rate` per second and
    the given uniqueness / reauth / insert-or-delete mix (insert and delete
    alternate)."""
    t, out, toggle = 0.0, [], 0
    kinds = rng.choice(3, size=n, p=list(mix))
    for k, kind in enumerate(kinds):
        t += rng.exponential(1.0 / rate)
        if kind == 0:
            r = make_uniqueness(k)
        elif kind == 1:
            r = make_reauth(k)
        else:
            r = make_insert(k) if toggle == 0 else make_delete(k)
            toggle ^= 1
        r.arrival, r.rid = t, k
        out.append(r)
    return out


def latency_stats(reqs, kind=None):
    sel = [r for r in reqs if kind is None or r.kind == kind]
    unfinished = [r for r in sel if r.done < 0 or r.error]
    if unfinished:
        raise RuntimeError(f"{len(unfinished)} request(s) unfinished or failed")
    lat = np.array([r.done - r.arrival for r in sel])
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3)}
