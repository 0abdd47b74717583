"""Serving form: a request loop with continuous batching over the DB scan.

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
- A request leaves when it has covered every chunk of the DB as it was when
  it joined, so requests that arrive during a scan start at the next step
  instead of waiting for the scan to finish.
- Every step runs one share GEMM per matrix and party for all active queries,
  then the reshare, the threshold circuits and the open into each request's
  bitmap row (slot map).
- Reauth runs on its own small engine set and streams, concurrently with the
  scan.
- Insertions and deletions are applied between steps.

The opened bits do not depend on the mix (the masks do), so each result
equals the result of the same request run alone (ServingEngines.run_alone,
the plain per-request form).
"""
import collections
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import engine as E
from . import synth

ROT = E.ROTATIONS
CHUNK = E.DB_CHUNK_SIZE


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
    covered: set = field(default_factory=set)
    slot: int = -1


def dummy_shares(rand_dataset=0x0de1e7e, index=0):
    """Per-party shares of the default iris (code all zeros, mask all ones)."""
    code = np.zeros((1, 200), np.uint64)
    mask = np.full((1, 200), np.uint64(0xFFFFFFFFFFFFFFFF))
    sh = synth.shares_of(code, mask, rand_dataset, index)
    return [(sh[p][0][0], sh[p][1][0]) for p in range(3)]


def bitmap_rows(words, n):
    """Set bits below n of a little-endian u64 bitmap, ascending: only the
    nonzero words are unpacked (a 2M-row bitmap row has a handful)."""
    w = np.ascontiguousarray(words).view(np.uint64)
    nz = np.flatnonzero(w)
    if nz.size == 0:
        return []
    bits = np.unpackbits(w[nz].view(np.uint8).reshape(-1, 8), axis=1, bitorder="little")
    i, j = np.nonzero(bits)
    rows = nz[i].astype(np.int64) * 64 + j
    return [int(r) for r in rows[rows < n]]


class PartySet:
    """Engines of the three parties for one query-set size: codes/masks ShareDB,
    Circuits and DistanceComparator per party, one exchange, three streams.
    Engine seeds are derive_seed(party seed, nonce_base + 0 / 1 / 5): every
    engine set of one process takes its own nonce_base, so no two sets draw the
    same keystream (the reference takes a fresh kdf nonce per engine,
    actor.rs:382-388)."""

    def __init__(self, max_rows, db_capacity, seeds, tag0, nonce_base=0):
        self.ex = E.Exchange(timeout_s=300)
        self.codes, self.masks, self.circ, self.cmp, self.streams = [], [], [], [], []
        self.nonce_base = nonce_base
        for p in range(3):
            so, sp = seeds[p]

            def seed(n):
                return E.derive_seed(so, nonce_base + n), E.derive_seed(sp, nonce_base + n)

            self.codes.append(E.ShareDB(p, CHUNK, max_rows, E.IRIS_CODE_LENGTH, seed(0), comm=self.ex, tag=tag0))
            self.masks.append(E.ShareDB(p, CHUNK, max_rows, E.MASK_CODE_LENGTH, seed(1), comm=self.ex,
                                        tag=tag0 + 1))
            self.circ.append(E.Circuits(p, max_rows * CHUNK, max_rows * CHUNK // 64, seed(5), self.ex,
                                        tag=tag0 + 2))
            self.cmp.append(E.DistanceComparator(max_rows, db_capacity))
            self.streams.append(E.stream_create())

    def parties(self, code_dbs, mask_dbs):
        return [E.PartyEngines(self.codes[p], self.masks[p], self.circ[p], self.cmp[p], code_dbs[p], mask_dbs[p])
                for p in range(3)]

    def sync(self):
        for s in self.streams:
            E.stream_synchronize(s)


class ServingEngines:
    """DB shares of the three parties plus the scan and reauth engine sets.
    dbs = (code_dbs, mask_dbs), per party, serves DB slices that already hold
    n_rows rows and room for `capacity` (they are written by insertions and
    deletions); otherwise the slices are allocated here."""

    def __init__(self, n_rows, capacity, max_active=64, seed_base=0, threshold_a=None, dbs=None):
        import torch

        self.torch = torch
        self.n_rows, self.capacity, self.max_active = n_rows, capacity, max_active
        self.A = E.THRESHOLD_ANON_STATS_A if threshold_a is None else threshold_a
        base = [np.full(8, seed_base + p, np.uint32) for p in range(3)]
        seeds = [(base[p], base[(p + 2) % 3]) for p in range(3)]
        self.scan = PartySet(max_active * ROT, capacity, seeds, 0, nonce_base=0)
        # reauth: one query, one chunk of 64 rows; its own keystreams
        self.small = PartySet(ROT, capacity, seeds, 8, nonce_base=16)
        if dbs is None:
            self.code_dbs = [self.scan.codes[p].alloc_db(capacity) for p in range(3)]
            self.mask_dbs = [self.scan.masks[p].alloc_db(capacity) for p in range(3)]
        else:
            self.code_dbs, self.mask_dbs = list(dbs[0]), list(dbs[1])
        self.scan_parties = self.scan.parties(self.code_dbs, self.mask_dbs)
        self.small_parties = self.small.parties(self.code_dbs, self.mask_dbs)
        words = -(-capacity // 64)
        self.words = words
        # one bitmap row of `capacity` bits per slot, per party; the small set's single row
        self.bitmaps = [torch.zeros(max_active * words + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
        self.small_bitmaps = [torch.zeros(words + 2, dtype=torch.int64, device="cuda") for _ in range(3)]
        self.slot_maps = [torch.zeros(max_active, dtype=torch.int32, device="cuda") for _ in range(3)]
        self.lock = threading.Lock()  # DB writes vs the reauth worker's reads
        self._dummy = None
        self._slots = None  # active slots the device slot maps hold
        self._inflight = collections.deque()  # per queued scan step: one event per party stream
        self._xstreams = [torch.cuda.ExternalStream(s) for s in self.scan.streams]
        self._small_xstreams = [torch.cuda.ExternalStream(s) for s in self.small.streams]
        self._small_qsets = None  # the reauth engines' query sets, refilled per request
        self.step_table = None  # {active-query bucket: chunks per step} from tune_steps
        self._alone_qsets = None  # the plain form's query sets, refilled per request
        self._readout = None  # pinned host copy of party 0's bitmap rows, one per slot
        self._pinned, self._pinned_ev, self._pin_next = None, [None] * 3, 0  # refill staging ring

    def sync_scan(self):
        """Waits for every queued scan step."""
        self.scan.sync()
        self._inflight.clear()

    # ------------------------------------------------------------------ DB
    def build(self, n_unique, dataset, rand_dataset):
        synth.build_db(self.code_dbs, self.mask_dbs, self.n_rows, n_unique, dataset, rand_dataset)

    def insert(self, shares):
        with self.lock:
            self.sync_scan()  # no queued step may read the rows being written
            row = self.n_rows
            if row >= self.capacity - CHUNK:
                raise ValueError(f"DB capacity exhausted ({row} rows of {self.capacity})")
            for p in range(3):
                self.code_dbs[p].load_records(row, shares[p][0].reshape(1, -1))
                self.mask_dbs[p].load_records(row, shares[p][1].reshape(1, -1))
            self.n_rows += 1
            return row

    def delete(self, row):
        if not 0 <= row < self.n_rows:
            raise ValueError(f"deletion of row {row} outside the DB ({self.n_rows} rows)")
        if self._dummy is None:
            self._dummy = dummy_shares()
        with self.lock:
            self.sync_scan()
            for p in range(3):
                self.code_dbs[p].load_records(row, self._dummy[p][0].reshape(1, -1))
                self.mask_dbs[p].load_records(row, self._dummy[p][1].reshape(1, -1))
        return row

    # ------------------------------------------------------------------ queries
    def _query_sets(self, ps, shares_list, rows, reuse=None):
        """The query sets of the three parties; `reuse` (earlier sets made with
        at least `rows` rows) is refilled in place instead of reallocated, once
        the queued steps that read them have finished."""
        if reuse is not None:
            if ps is self.scan:
                self.sync_scan()
            else:
                ps.sync()
        qc, qm = [], []
        for p in range(3):
            code = np.stack([s[p][0] for s in shares_list])
            mask = np.stack([s[p][1] for s in shares_list])
            if reuse is None:
                qc.append(ps.codes[p].upload_query_shares_rows(code, rows))
                qm.append(ps.masks[p].upload_query_shares_rows(mask, rows))
            else:
                qc.append(ps.codes[p].refill_query_shares(reuse[0][p], code, rows))
                qm.append(ps.masks[p].refill_query_shares(reuse[1][p], mask, rows))
        return qc, qm

    def refill_scan(self, qsets, shares_list, rows):
        """Refills the scan engines' query sets (made at full size) with
        shares_list on the parties' scan streams: ordered after the queued
        steps that still read the old sets and before the steps queued later,
        with no synchronisation. The shares pass through pinned host buffers,
        a ring of three, each reused once the refill that read it has run."""
        torch = self.torch
        if self._pinned is None:
            self._pinned = [[[torch.empty((self.max_active, K), dtype=torch.int16, pin_memory=True)
                              for K in (E.IRIS_CODE_LENGTH, E.MASK_CODE_LENGTH)] for _ in range(3)]
                            for _ in range(3)]
        k = self._pin_next
        self._pin_next = (k + 1) % 3
        for ev in self._pinned_ev[k] or ():
            ev.synchronize()
        n, evs = len(shares_list), []
        for p in range(3):
            bc = self._pinned[k][p][0].numpy().view(np.uint16)
            bm = self._pinned[k][p][1].numpy().view(np.uint16)
            for j, sh in enumerate(shares_list):
                bc[j] = sh[p][0]
                bm[j] = sh[p][1]
            st = self.scan.streams[p]
            self.scan.codes[p].refill_query_shares_async(qsets[0][p], bc.ctypes.data, n, rows, st)
            self.scan.masks[p].refill_query_shares_async(qsets[1][p], bm.ctypes.data, n, rows, st)
            ev = torch.cuda.Event()
            ev.record(self._xstreams[p])
            evs.append(ev)
        self._pinned_ev[k] = evs
        return qsets

    def n_chunks(self, rows):
        return max(1, -(-rows // CHUNK))

    def chunks_per_step(self, batch, chunk_index, n_chunks):
        """Chunks one step covers: several while few queries are active (their
        launches alone would leave the GPU partly idle), one under load so that
        arrivals join at short intervals; bounded by the result buffers
        (chunks x rows within the engines' query rows) and the ring's end.
        With a table from tune_steps, the entry of the largest bucket at or
        below `batch`; otherwise a fixed rule."""
        if self.step_table:
            under = [b for b in sorted(self.step_table) if b <= batch]
            k = self.step_table[under[-1] if under else min(self.step_table)]
        else:
            k = 8 if batch == 1 else 4 if batch <= 2 else 2 if batch <= 8 else 1
        k = min(k, max(1, (self.max_active * ROT) // (batch * ROT)), n_chunks - chunk_index)
        return max(1, k)

    def time_step(self, batch, n_chunks, reps=2):
        """Seconds of one scan step of `batch` queries (default-iris shares)
        over n_chunks chunks from the ring's start, best of `reps` after one
        warm-up; the slots it writes are cleared when requests take them."""
        if self._dummy is None:
            self._dummy = dummy_shares()
        qc, qm = self._query_sets(self.scan, [self._dummy] * batch, batch * ROT)
        slots = list(range(batch))
        best = float("inf")
        for r in range(reps + 1):
            self.sync_scan()
            t0 = time.perf_counter()
            self.step(qc, qm, batch, 0, self.n_rows, slots, n_chunks=n_chunks)
            self.sync_scan()
            if r:
                best = min(best, time.perf_counter() - t0)
        return best

    def tune_steps(self, tuner=None, **kw):
        """Step sizes per active-query count from the autotuner (batching.Autotuner)."""
        from .batching import Autotuner

        self.step_table = (tuner or Autotuner()).step_chunks(self, **kw)
        return self.step_table

    def step(self, qc, qm, batch, chunk_index, db_rows, slots, max_queued=2, n_chunks=1):
        """Queues one step of the scan engines for the active query set over
        chunks [chunk_index, chunk_index + n_chunks). At most `max_queued` steps
        wait on the scan streams, so the next step is queued while one runs; the
        device slot maps are rewritten only when the active slots change, after
        the steps that read them."""
        slots = list(slots)
        if slots != self._slots:
            # each party's slot map is rewritten on its scan stream: after the
            # queued steps that read the old map, before the step below
            host = self.torch.tensor(slots, dtype=self.torch.int32).pin_memory()
            for p in range(3):
                with self.torch.cuda.stream(self._xstreams[p]):
                    self.slot_maps[p][:len(slots)].copy_(host, non_blocking=True)
            self._slots = slots
        while len(self._inflight) >= max_queued:
            for ev in self._inflight.popleft():
                ev.synchronize()
        E.scan_chunks_trio(self.scan_parties, qc, qm, chunk_index * CHUNK, n_chunks, db_rows, batch, self.A,
                           [b.data_ptr() for b in self.bitmaps], self.scan.streams,
                           [m.data_ptr() for m in self.slot_maps], row_bits=self.words * 64)
        evs = []
        for xs in self._xstreams:
            ev = self.torch.cuda.Event()
            ev.record(xs)
            evs.append(ev)
        self._inflight.append(evs)

    def clear_slot(self, slot):
        """Zeroes a slot's bitmap rows with fill kernels on the scan streams:
        after the queued steps (and readouts) of the slot's previous request,
        before every step queued later; nothing waits."""
        for b, s in zip(self.bitmaps, self.scan.streams):
            E.fill_u64(b.data_ptr() + 8 * slot * self.words, self.words, 0, s)

    def begin_readout(self, slot):
        """Queues the copy of party 0's bitmap row of `slot` into pinned host
        memory on party 0's scan stream, after the steps queued so far, and
        returns the event that marks its completion (nothing waits)."""
        if self._readout is None:
            self._readout = self.torch.empty((self.max_active, self.words), dtype=self.torch.int64,
                                             pin_memory=True)
        xs = self._xstreams[0]
        with self.torch.cuda.stream(xs):
            self._readout[slot].copy_(self.bitmaps[0][slot * self.words:(slot + 1) * self.words],
                                      non_blocking=True)
            ev = self.torch.cuda.Event()
            ev.record(xs)
        return ev

    def readout_rows(self, slot, db_rows):
        """The rows below db_rows set in the slot's copied bitmap row (after its
        begin_readout event completed)."""
        return bitmap_rows(self._readout[slot].numpy(), db_rows)

    def slot_matches(self, slot, db_rows):
        self.sync_scan()
        return bitmap_rows(self.bitmaps[0][slot * self.words:(slot + 1) * self.words].cpu().numpy(), db_rows)

    def reauth(self, shares, target):
        """One query x 31 rotations against DB row `target` on the small engine set."""
        if not 0 <= target < self.n_rows:
            raise ValueError(f"reauth target {target} outside the DB ({self.n_rows} rows)")
        # refilled in place: an allocation or free would synchronise the whole
        # device and wait for the scan steps queued by the request loop
        if self._small_qsets is None:
            self._small_qsets = self._query_sets(self.small, [shares], ROT)
        qc, qm = self._query_sets(self.small, [shares], ROT, reuse=self._small_qsets)
        off = target // 64 * 64
        for b, xs in zip(self.small_bitmaps, self._small_xstreams):
            with self.torch.cuda.stream(xs):  # ordered before this party's scan step
                b.zero_()
        with self.lock:  # the DB rows are not written meanwhile
            # a 64-row window: the scan chunk of [off, off + 64) of a DB ending at off + 64
            E.scan_chunk_trio(self.small_parties, qc, qm, off, off + 64, 1, self.A,
                              [b.data_ptr() for b in self.small_bitmaps], self.small.streams)
            self.small.sync()
        bits = self.small_bitmaps[0][: (off + 64) // 64 + 1].cpu().numpy().view(np.uint8)
        return bool(np.unpackbits(bits, bitorder="little")[target])

    def run_alone(self, shares):
        """The plain per-request form: a full scan of one uniqueness query."""
        # one persistent one-query set, refilled per request (no allocation)
        if self._alone_qsets is None:
            self._alone_qsets = self._query_sets(self.scan, [shares], ROT)
        qc, qm = self._query_sets(self.scan, [shares], ROT, reuse=self._alone_qsets)
        rows = self.n_rows
        self.clear_slot(0)
        c, n = 0, self.n_chunks(rows)
        while c < n:
            k = self.chunks_per_step(1, c, n)
            self.step(qc, qm, 1, c, rows, [0], n_chunks=k)
            c += k
        return self.slot_matches(0, rows)


class Server:
    """Real-time request loop over ServingEngines (one scheduler thread plus a
    reauth worker). Admission: a uniqueness request takes a free slot while
    free HBM stays above `reserve_bytes` plus its device footprint."""

    def __init__(self, eng, reserve_bytes=4 << 30, max_active=None):
        self.eng = eng
        self.max_active = max_active or eng.max_active
        self.reserve = reserve_bytes
        self.footprint = 3 * ROT * (E.IRIS_CODE_LENGTH + E.MASK_CODE_LENGTH) * 2 * 2  # limbs + staging
        self.reauth_q = queue.Queue()
        self.done = []
        self.steps = 0
        self.admission_waits = 0
        self.max_concurrent = 0
        self.failed = []  # requests whose serving raised (run() raises after the loop)

    def _free_hbm(self):
        free, _ = self.eng.torch.cuda.mem_get_info()
        return free

    def _collect(self, reading, free_slots):
        """Finishes the requests whose bitmap readout has completed (in order)
        and frees their slots."""
        while reading and reading[0][1].query():
            r, _ = reading.pop(0)
            r.result = self.eng.readout_rows(r.slot, r.db_rows)
            r.done = time.perf_counter() - self.t0
            free_slots.append(r.slot)
            self.done.append(r)

    def _reauth_worker(self):
        while True:
            r = self.reauth_q.get()
            if r is None:
                return
            try:
                r.result = self.eng.reauth(r.shares, r.target)
            except Exception as e:  # noqa: BLE001 - recorded on the request, raised by run()
                r.error = f"{type(e).__name__}: {e}"
                self.failed.append(r)
            r.db_rows = self.eng.n_rows
            r.done = time.perf_counter() - self.t0
            self.done.append(r)

    def run(self, requests):
        """Serves `requests` (sorted by arrival, seconds from now); returns them
        with results and completion times."""
        eng = self.eng
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        worker = threading.Thread(target=self._reauth_worker, daemon=True)
        worker.start()
        waiting, active, reading = [], [], []
        free_slots = list(range(self.max_active))
        qsets, qsets_full, active_ids = None, None, None
        i, cursor = 0, 0
        while i < len(reqs) or waiting or active or reading:
            self._collect(reading, free_slots)
            now = time.perf_counter() - self.t0
            while i < len(reqs) and reqs[i].arrival <= now:
                r = reqs[i]
                i += 1
                if r.kind == "uniqueness":
                    waiting.append(r)
                elif r.kind == "reauth":
                    self.reauth_q.put(r)
                else:
                    try:
                        r.result = eng.insert(r.shares) if r.kind == "insert" else eng.delete(r.target)
                    except ValueError as e:
                        r.error = str(e)
                        self.failed.append(r)
                    r.db_rows = eng.n_rows
                    r.done = time.perf_counter() - self.t0
                    self.done.append(r)
            # admission between chunk steps
            while waiting and free_slots:
                # with nothing active one request always goes in (progress)
                if active and self._free_hbm() < self.reserve + self.footprint * (len(active) + 1):
                    self.admission_waits += 1
                    break
                r = waiting.pop(0)
                r.slot = free_slots.pop(0)
                r.db_rows = eng.n_rows
                r.covered = set()
                eng.clear_slot(r.slot)
                active.append(r)
            self.max_concurrent = max(self.max_concurrent, len(active))
            if not active:
                if reading and not waiting:
                    reading[0][1].synchronize()  # nothing to scan: wait for the oldest readout
                elif i < len(reqs):
                    time.sleep(max(0.0, min(0.002, reqs[i].arrival - (time.perf_counter() - self.t0))))
                continue
            # the active set's identity: request objects, not their (caller-set) rids
            ids = [id(r) for r in active]
            if ids != active_ids:
                # one set per party and matrix, made once at full size and refilled
                if qsets_full is None:
                    first = [r.shares for r in active]
                    pad = first + [first[0]] * (self.max_active - len(first))
                    qsets_full = eng._query_sets(eng.scan, pad, self.max_active * ROT)
                qsets = eng.refill_scan(qsets_full, [r.shares for r in active], len(active) * ROT)
                active_ids = ids
            n_chunks = eng.n_chunks(eng.n_rows)
            cursor %= n_chunks
            # no lock: scan steps only read the DB, and the DB writes (insert,
            # delete) run on this thread after the queued steps finished
            k = eng.chunks_per_step(len(active), cursor, n_chunks)
            eng.step(qsets[0], qsets[1], len(active), cursor, eng.n_rows, [r.slot for r in active],
                     n_chunks=k)
            self.steps += 1
            for r in active:
                r.covered.update(range(cursor, cursor + k))
            cursor += k
            finished = [r for r in active if len(r.covered) >= eng.n_chunks(r.db_rows)
                        and all(c in r.covered for c in range(eng.n_chunks(r.db_rows)))]
            for r in finished:  # its bitmap row is copied out after its last step
                reading.append((r, eng.begin_readout(r.slot)))
                active.remove(r)
            self._collect(reading, free_slots)
        self.reauth_q.put(None)
        worker.join()
        if self.failed:
            r = self.failed[0]
            raise RuntimeError(f"{len(self.failed)} request(s) failed; first: {r.kind} rid {r.rid}: {r.error}")
        return reqs


class PerBatchServer:
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
        self.eng, self.max_batch = eng, max_batch
        self.batches = 0
        self.failed = []

    def run(self, requests):
        eng, B = self.eng, self.max_batch
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        if eng._dummy is None:
            eng._dummy = dummy_shares()
        qsets = None
        i, pending = 0, []
        while i < len(reqs) or pending:
            now = time.perf_counter() - self.t0
            while i < len(reqs) and reqs[i].arrival <= now:
                pending.append(reqs[i])
                i += 1
            if not pending:
                time.sleep(max(0.0, min(0.002, reqs[i].arrival - now)))
                continue
            batch = []
            while pending and len(batch) < B:
                r = pending.pop(0)
                if r.kind in ("insert", "delete"):
                    try:
                        r.result = eng.insert(r.shares) if r.kind == "insert" else eng.delete(r.target)
                    except ValueError as e:
                        r.error = str(e)
                        self.failed.append(r)
                    r.db_rows = eng.n_rows
                    r.done = time.perf_counter() - self.t0
                else:
                    batch.append(r)
            if not batch:
                continue
            shares = [r.shares for r in batch] + [eng._dummy] * (B - len(batch))
            if qsets is None:
                qsets = eng._query_sets(eng.scan, shares, B * ROT)
            qsets = eng._query_sets(eng.scan, shares, B * ROT, reuse=qsets)
            rows = eng.n_rows
            slots = list(range(len(batch)))
            for s in slots:
                eng.clear_slot(s)
            n = eng.n_chunks(rows)
            for c in range(n):  # the reference's chunk loop: one chunk per step
                eng.step(qsets[0], qsets[1], len(batch), c, rows, slots, n_chunks=1)
            self.batches += 1
            for s, r in enumerate(batch):
                matches = eng.slot_matches(s, rows)
                if r.kind == "reauth":
                    if 0 <= r.target < rows:
                        r.result = r.target in set(matches)
                    else:
                        r.error = f"reauth target {r.target} outside the DB ({rows} rows)"
                        self.failed.append(r)
                else:
                    r.result = matches
                r.db_rows = rows
                r.done = time.perf_counter() - self.t0
        if self.failed:
            r = self.failed[0]
            raise RuntimeError(f"{len(self.failed)} request(s) failed; first: {r.kind} rid {r.rid}: {r.error}")
        return reqs


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
