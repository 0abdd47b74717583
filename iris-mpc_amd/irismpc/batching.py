"""Batching form: query batches of uneven size packed into few launches.

The plain form runs one batch (1-256 queries x 31 rotations) at a time: for
every DB chunk one share-GEMM launch per matrix. A batch of a few queries
reads the whole DB share for little MFMA work (the GEMM is HBM-bound below
~7 queries), and the 4-query tiles of the rotation-aware kernel are only
partly used for small batches.

The batching form concatenates consecutive batches' query shares into one
packed query set (up to `max_queries` queries, at most 64 batches). One launch
per chunk and matrix then serves all of them. Each batch keeps the output
layout and ChaCha12 mask counters of its own plain run (segment table of the
packed launch), so both forms produce identical shares.

The whole packed schedule is captured once into a HIP graph and replayed.
The Autotuner times the GEMM variants and the pack size on the device and
caches its choices.
"""
import ctypes
import json
import math
import os
import time

import numpy as np

from ._lib import call
from .engine import Graph, gemm_rot_variants, stream_synchronize

ROTATIONS = 31


def log_uniform_batch_sizes(n, lo=1, hi=256, seed=0):
    """n batch sizes from a seeded log-uniform distribution over [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.clip(np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n))), lo, hi).astype(int)


def _packs(pack_of, n_packs):
    packs = [[] for _ in range(n_packs)]
    for i, p in enumerate(pack_of):
        packs[int(p)].append(i)
    return packs


def plan_packs(sizes, max_queries, max_batches=64):
    """Greedy packing of consecutive batches (the library's planner): a pack
    holds at most max_queries queries (a larger batch gets a pack of its own)
    and at most max_batches. Returns lists of batch indices."""
    sz = np.ascontiguousarray(sizes, dtype=np.uint32)
    pack_of = np.zeros(max(1, sz.size), np.uint32)
    n = ctypes.c_size_t()
    call("irismpc_batch_plan_packs", sz.ctypes.data, sz.size, max_queries, max_batches, pack_of.ctypes.data,
         ctypes.byref(n))
    return _packs(pack_of[:sz.size], n.value)


def counter_advance(total, chunk, rows):
    """ChaCha blocks a remote engine's dot_reduce_range uses over `total` rows
    (mask + reshare OTP per chunk)."""
    v = ctypes.c_uint64()
    call("irismpc_batch_counter_advance", total, chunk, rows, ctypes.byref(v))
    return v.value


def range_advance(n_rows, chunk, rows):
    """ChaCha blocks of a whole scan of n_rows DB rows, launched one chunk at a
    time with the last chunk padded to 64 rows (as the runners launch it)."""
    v = ctypes.c_uint64()
    call("irismpc_batch_scan_counter_advance", n_rows, chunk, rows, ctypes.byref(v))
    return v.value


def pick_step_chunks(per_chunk, tol=0.03, max_step=None):
    """Chunks per scan step from measured seconds per chunk {k: s} (the
    library's rule): steps longer than max_step seconds are left out (when
    every step is longer, the smallest k); of the rest, the smallest k within
    `tol` of the fastest per chunk."""
    ks = sorted(per_chunk)
    k = (ctypes.c_uint32 * len(ks))(*ks)
    t = (ctypes.c_double * len(ks))(*[per_chunk[x] for x in ks])
    out = ctypes.c_uint32()
    call("irismpc_serving_pick_step", k, t, len(ks), float(tol), float(max_step or 0.0), ctypes.byref(out))
    return out.value


class Matrix:
    """One matrix (code or mask) of one party: engine, DB share, multiplier and
    the raw query shares of every batch ([b][K] u16 on the host)."""

    def __init__(self, engine, db, n_rows, multiplier, shares):
        self.engine, self.db, self.n_rows, self.mult, self.shares = engine, db, n_rows, multiplier, shares


class BatchRunner:
    """Plain and packed runs of the same batches over one or more matrices,
    one native runner (irismpc_batch_runner_*) per matrix; each batch's last
    chunk of results stays in its own output region."""

    def __init__(self, matrices, sizes, chunk, stream, max_queries=256, max_batches=64):
        import torch

        self.m, self.sizes, self.chunk, self.stream = matrices, [int(b) for b in sizes], chunk, stream
        self.outs = [[torch.zeros(chunk * b * ROTATIONS + 64, dtype=torch.int16, device="cuda")
                      for b in self.sizes] for _ in matrices]
        torch.cuda.synchronize()  # zeroed before the library's stream writes them
        sz = np.asarray(self.sizes, np.uint32)
        self.h = []
        for k, m in enumerate(matrices):
            shares = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint16) for x in m.shares]))
            outs = (ctypes.c_void_p * len(self.sizes))(*[o.data_ptr() for o in self.outs[k]])
            h = ctypes.c_void_p()
            call("irismpc_batch_runner_create", m.engine.h, m.db.h, m.n_rows, m.mult, sz.ctypes.data, sz.size,
                 shares.ctypes.data, outs, chunk, max_queries, max_batches, stream, ctypes.byref(h))
            self.h.append(h.value)
        self.max_queries, self.graph, self.graph_persistent, self.prepared = max_queries, None, False, None
        self.plan = self._plan()

    def __del__(self):
        for h in getattr(self, "h", []):
            call("irismpc_batch_runner_destroy", h)
        self.h = []

    def _plan(self):
        pack_of, n = np.zeros(len(self.sizes), np.uint32), ctypes.c_size_t()
        call("irismpc_batch_runner_plan", self.h[0], pack_of.ctypes.data, ctypes.byref(n))
        return _packs(pack_of, n.value)

    def replan(self, max_queries):
        """New pack size: rebuilds the plan and the packed query sets (the
        output regions and the plain query sets stay)."""
        self.graph, self.prepared = None, None
        for h in self.h:
            call("irismpc_batch_runner_replan", h, max_queries, self.stream)
        self.max_queries, self.plan = max_queries, self._plan()

    def comparisons(self):
        return sum(b * ROTATIONS for b in self.sizes) * self.m[0].n_rows

    def run_plain(self, ctr0=0):
        """One batch at a time, each to completion (stream synchronised)."""
        for m in self.m:
            m.engine.counter = ctr0
        for i in range(len(self.sizes)):
            for h in self.h:
                call("irismpc_batch_runner_issue_plain", h, i, self.stream)
            stream_synchronize(self.stream)

    def _issue_packed(self, ctr0, n_rows=None, persistent=False):
        for h in self.h:
            if persistent:
                call("irismpc_batch_runner_issue_persistent", h, self.stream)
            else:
                call("irismpc_batch_runner_issue_packed", h, ctr0, n_rows or 0, self.stream)

    def _prepare(self, ctr0, persistent):
        if persistent and self.prepared != ctr0:
            for h in self.h:
                call("irismpc_batch_runner_prepare_persistent", h, ctr0, self.stream)
            self.prepared = ctr0

    def capture(self, ctr0=0, persistent=False):
        """Captures the packed schedule into a HIP graph (persistent: one
        launch per chunk and matrix covering every pack)."""
        self._prepare(ctr0, persistent)
        self.graph = Graph(self.stream).capture()
        try:
            self._issue_packed(ctr0, persistent=persistent)
        finally:
            self.graph.end()
        self.graph_persistent = persistent
        return self.graph

    def run_packed(self, ctr0=0, graph=True, n_rows=None, persistent=False):
        """The packed schedule (replayed from its graph); n_rows limits the scan
        to the first n_rows DB rows (no graph, one launch per pack)."""
        if graph and n_rows is None:
            if self.graph is None or self.graph_persistent != persistent:
                self.capture(ctr0, persistent)
            self.graph.launch()
        else:
            self._prepare(ctr0, persistent and n_rows is None)
            self._issue_packed(ctr0, n_rows, persistent=persistent and n_rows is None)
        stream_synchronize(self.stream)


class Autotuner:
    """Picks the GEMM variant for a matrix and the pack size of a plan, and
    the serving form's step size (chunks per threshold comparison) per number
    of active queries, by timing them on the device; choices are cached per
    device and shape in a JSON file (default iris-mpc_amd/_build/autotune.json)."""

    def __init__(self, path=None):
        here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        self.path = path or os.environ.get("IRISMPC_AUTOTUNE", os.path.join(here, "_build", "autotune.json"))
        try:
            with open(self.path) as f:
                self.cache = json.load(f)
        except (OSError, ValueError):
            self.cache = {}
        self.log = []

    def _save(self):
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        with open(self.path, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)

    @staticmethod
    def _device():
        import torch

        return torch.cuda.get_device_name(0).replace(" ", "_")

    def _time(self, fn, stream, reps=2):
        from .engine import stream_synchronize

        fn()
        stream_synchronize(stream)
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            stream_synchronize(stream)
            best = min(best, time.perf_counter() - t0)
        return best

    def gemm_variant(self, m, query, rows, chunk, stream, sample_rows=4 * 32768, candidates=None):
        """Fastest variant for query sets of `rows` rows on matrix m (raw-share queries)."""
        import torch

        key = f"{self._device()}/gemm/K={m.engine.code_length}/rows={rows}"
        if key in self.cache:
            return self.cache[key]["variant"]
        cands = candidates or gemm_rot_variants()
        n = min(sample_rows, m.n_rows)
        out = torch.empty(n * rows + 64, dtype=torch.int16, device="cuda")
        times = {}
        keep = m.engine.counter
        for v in cands:
            m.engine.set_variant(v)
            times[v] = self._time(lambda: m.engine.dot_reduce_range(query, m.db, 0, -(-n // 64) * 64, chunk,
                                                                     m.mult, out.data_ptr(), stream), stream)
        m.engine.counter = keep
        best = min(times, key=times.get)
        m.engine.set_variant(best)
        self.cache[key] = {"variant": best, "seconds": {str(k): v for k, v in times.items()}}
        self.log.append((key, times))
        self._save()
        return best

    def pack_size(self, runner, sample_rows, candidates=(64, 128, 256)):
        """Pack size with the best measured time for the runner's batches over
        the first sample_rows DB rows; the runner is left on that pack size."""
        sizes = runner.sizes
        key = f"{self._device()}/pack/n={len(sizes)}/q={int(np.sum(sizes))}/c={','.join(map(str, candidates))}"
        if key in self.cache:
            best = self.cache[key]["max_queries"]
            if runner.max_queries != best:
                runner.replan(best)
            return best
        times = {}
        for mq in candidates:
            runner.replan(mq)
            times[mq] = self._time(lambda: runner.run_packed(graph=False, n_rows=sample_rows), runner.stream,
                                   reps=1)
        best = min(times, key=times.get)
        if runner.max_queries != best:
            runner.replan(best)
        self.cache[key] = {"max_queries": best, "seconds": {str(k): v for k, v in times.items()}}
        self.log.append((key, times))
        self._save()
        return best

    def step_chunks(self, eng, buckets=(1, 2, 4, 8, 16, 32, 64), ks=(1, 2, 4, 8, 16), tol=0.03,
                    max_step_s=0.008):
        """Chunks per scan step of the serving form (`eng`: ServingEngines) for
        each bucket of active queries: the library times one step of that many
        queries for every k that the result buffers and the DB allow and picks
        (pick_step_chunks; steps of at most max_step_s seconds). A bucket's
        choice serves active counts from it up to the next bucket. Returns
        {bucket: k}."""
        key = (f"{self._device()}/steps/rows={eng.n_rows}/max_active={eng.max_active}"
               f"/b={','.join(map(str, buckets))}/k={','.join(map(str, ks))}/tol={tol}/max={max_step_s}")
        if key in self.cache:
            return {int(b): k for b, k in self.cache[key]["chunks"].items()}
        nb, nk = len(buckets), len(ks)
        b = (ctypes.c_uint32 * nb)(*buckets)
        k = (ctypes.c_uint32 * nk)(*ks)
        chunks, secs = (ctypes.c_uint32 * nb)(), (ctypes.c_double * (nb * nk))()
        call("irismpc_serving_tune_steps", eng.h, b, nb, k, nk, float(tol), float(max_step_s), chunks, secs)
        table = {int(buckets[i]): int(chunks[i]) for i in range(nb) if chunks[i]}
        times = {str(buckets[i]): {str(ks[j]): secs[i * nk + j] for j in range(nk) if secs[i * nk + j] >= 0}
                 for i in range(nb) if chunks[i]}
        self.cache[key] = {"chunks": {str(x): y for x, y in table.items()}, "seconds_per_chunk": times}
        self.log.append((key, times))
        self._save()
        return table
