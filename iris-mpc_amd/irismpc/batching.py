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
import json
import math
import os
import time

import numpy as np

from .engine import Graph, gemm_rot_variants

ROTATIONS = 31


def log_uniform_batch_sizes(n, lo=1, hi=256, seed=0):
    """n batch sizes from a seeded log-uniform distribution over [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.clip(np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n))), lo, hi).astype(int)


def plan_packs(sizes, max_queries, max_batches=64):
    """Greedy packing of consecutive batches: a pack holds at most max_queries
    queries (a larger batch gets a pack of its own) and at most max_batches."""
    packs, cur, q = [], [], 0
    for i, b in enumerate(sizes):
        if cur and (q + b > max_queries or len(cur) == max_batches):
            packs.append(cur)
            cur, q = [], 0
        cur.append(i)
        q += int(b)
    if cur:
        packs.append(cur)
    return packs


def counter_advance(total, chunk, rows):
    """ChaCha blocks a remote engine's dot_reduce_range uses over `total` rows
    (mask + reshare OTP per chunk)."""
    n, c0 = 0, 0
    while c0 < total:
        n += 2 * -(-(min(chunk, total - c0) * rows * 2) // 64)
        c0 += chunk
    return n


def range_advance(n_rows, chunk, rows):
    """ChaCha blocks of a whole scan of n_rows DB rows, launched one chunk at a
    time with the last chunk padded to 64 rows (as the runners launch it)."""
    return sum(counter_advance(tot, chunk, rows) for _, tot in _ranges(n_rows, chunk))


class Matrix:
    """One matrix (code or mask) of one party: engine, DB share, multiplier and
    the raw query shares of every batch ([b][K] u16 on the host)."""

    def __init__(self, engine, db, n_rows, multiplier, shares):
        self.engine, self.db, self.n_rows, self.mult, self.shares = engine, db, n_rows, multiplier, shares


def _ranges(n_rows, chunk):
    off = 0
    while off < n_rows:
        yield off, -(-min(chunk, n_rows - off) // 64) * 64
        off += chunk


class BatchRunner:
    """Plain and packed runs of the same batches over one or more matrices;
    each batch's last chunk of results stays in its own output region."""

    def __init__(self, matrices, sizes, chunk, stream, max_queries=256, max_batches=64):
        import torch

        self.m, self.sizes, self.chunk, self.stream = matrices, [int(b) for b in sizes], chunk, stream
        self.plan = plan_packs(self.sizes, max_queries, max_batches)
        self.outs = [[torch.zeros(chunk * b * ROTATIONS + 64, dtype=torch.int16, device="cuda")
                      for b in self.sizes] for _ in matrices]
        self.q_plain = [[m.engine.upload_query_shares_rows(m.shares[i], b * ROTATIONS)
                         for i, b in enumerate(self.sizes)] for m in matrices]
        self.max_batches = max_batches
        self.q_packed = []
        self.replan(max_queries)

    def replan(self, max_queries):
        """New pack size: rebuilds the plan and the packed query sets (the
        output regions and the plain query sets stay)."""
        self.plan = plan_packs(self.sizes, max_queries, self.max_batches)
        self.max_queries = max_queries
        self.q_packed, self.graph = [], None  # old sets are freed before the new ones exist
        for m in self.m:
            sets = []
            for pack in self.plan:
                raw = np.concatenate([m.shares[i] for i in pack])
                sets.append(m.engine.upload_query_shares_rows(raw, raw.shape[0] * ROTATIONS))
            self.q_packed.append(sets)

    def comparisons(self):
        return sum(b * ROTATIONS for b in self.sizes) * self.m[0].n_rows

    def run_plain(self, ctr0=0):
        """One batch at a time, each to completion (stream synchronised)."""
        from .engine import stream_synchronize

        for k, m in enumerate(self.m):
            m.engine.counter = ctr0
        for i in range(len(self.sizes)):
            for k, m in enumerate(self.m):
                for off, tot in _ranges(m.n_rows, self.chunk):
                    m.engine.dot_reduce_range(self.q_plain[k][i], m.db, off, tot, self.chunk, m.mult,
                                              self.outs[k][i].data_ptr(), self.stream)
            stream_synchronize(self.stream)

    def _issue_packed(self, ctr0, n_rows=None):
        for k, m in enumerate(self.m):
            n = m.n_rows if n_rows is None else n_rows
            base, c = [], ctr0
            for b in self.sizes:  # the plain form's counter at the start of each batch
                base.append(c)
                if m.engine.is_remote:
                    c += range_advance(n, self.chunk, b * ROTATIONS)
            for p, pack in enumerate(self.plan):
                rows = [self.sizes[i] * ROTATIONS for i in pack]
                outs = [self.outs[k][i].data_ptr() for i in pack]
                for off, tot in _ranges(n, self.chunk):
                    ctr = [base[i] + (counter_advance(off, self.chunk, r) if m.engine.is_remote else 0)
                           for i, r in zip(pack, rows)]
                    m.engine.dot_reduce_packed(self.q_packed[k][p], m.db, off, tot, self.chunk, m.mult,
                                               rows, outs, ctr, self.stream)

    def capture(self, ctr0=0):
        """Captures the packed schedule into a HIP graph."""
        self.graph = Graph(self.stream).capture()
        try:
            self._issue_packed(ctr0)
        finally:
            self.graph.end()
        return self.graph

    def run_packed(self, ctr0=0, graph=True, n_rows=None):
        """The packed schedule (replayed from its graph); n_rows limits the scan
        to the first n_rows DB rows (no graph)."""
        from .engine import stream_synchronize

        if graph and n_rows is None:
            if self.graph is None:
                self.capture(ctr0)
            self.graph.launch()
        else:
            self._issue_packed(ctr0, n_rows)
        stream_synchronize(self.stream)


def pick_step_chunks(per_chunk, tol=0.03, max_step=None):
    """Chunks per scan step from measured seconds per chunk {k: s}. Steps
    longer than max_step seconds are left out (a request arriving meanwhile
    waits up to a whole step to join the active set; when every step is
    longer, the smallest k). Of the rest, the smallest k within `tol` of the
    fastest per chunk: a larger step pays only by what it saves per chunk."""
    fits = {k: t for k, t in per_chunk.items() if max_step is None or k * t <= max_step}
    if not fits:
        return min(per_chunk)
    best = min(fits.values())
    return min(k for k, t in fits.items() if t <= best * (1 + tol))


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
        each bucket of active queries: every k that the result buffers and the
        DB allow is timed as one step of that many queries, and
        pick_step_chunks chooses (steps of at most max_step_s seconds). A
        bucket's choice serves active counts from it up to the next bucket.
        Returns {bucket: k}."""
        n_chunks = eng.n_chunks(eng.n_rows)
        key = (f"{self._device()}/steps/rows={eng.n_rows}/max_active={eng.max_active}"
               f"/b={','.join(map(str, buckets))}/k={','.join(map(str, ks))}/tol={tol}/max={max_step_s}")
        if key in self.cache:
            return {int(b): k for b, k in self.cache[key]["chunks"].items()}
        table, times = {}, {}
        for b in buckets:
            if b > eng.max_active:
                break
            cands = [k for k in ks if k * b <= eng.max_active and k <= n_chunks] or [1]
            per = {k: eng.time_step(b, k) / k for k in cands}
            table[b] = pick_step_chunks(per, tol, max_step_s)
            times[str(b)] = {str(k): v for k, v in per.items()}
        self.cache[key] = {"chunks": {str(b): k for b, k in table.items()}, "seconds_per_chunk": times}
        self.log.append((key, times))
        self._save()
        return table
