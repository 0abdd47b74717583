"""The reference server's two-stage matching flow for one batch, three parties
on one GPU (iris-mpc-gpu/src/server/actor.rs:860-1100):

1. the full-scan eye is compared against the whole HBM-resident DB at the
   prefilter threshold 0.375 (THRESHOLD_ANON_STATS_A, actor.rs:2246-2589);
2. the DB ids set in any query row of that bitmap, plus the batch's OR-rule
   and reauth targets, form the candidate subset (get_partial_results,
   distance_comparator.rs:545-648; actor.rs:938-970);
3. both eyes are compared against the subset only, at the matching threshold
   0.345 (THRESHOLD_A), and the opened bits land at the candidates' DB ids
   (compare_query_against_db_subset_and_self, actor.rs:1900-2244);
4. the two eye bitmaps are merged with the AND rule, or per (query, DB id) the
   OR rule (join_db_matches[_with_bitmaps], actor.rs:1097-1131).

With self_compare, each eye's queries are also compared against the batch's
own irises before its subset scan (compare_query_against_self,
actor.rs:1789-1897), and the batch bitmaps are merged after the DB ones
(join_batch_matches, actor.rs:1132-1137): a query matching an earlier query of
the batch in both eyes gets final result 0xFFFFFFFE and match id
0xFFFFFFFF - (that query's position).

Matches at 0.345 are a subset of the matches at 0.375, so the answer equals a
full scan of both eyes at 0.345 merged the same way, at the cost of one full
scan instead of two.
"""
import numpy as np

from .engine import (DB_CHUNK_SIZE, THRESHOLD_A, THRESHOLD_ANON_STATS_A, download,
                     scan_batch_self_trio, scan_db_trio, scan_subset_trio, stream_create,
                     stream_synchronize)

EYES = ("left", "right")


def prepare_or_policy_bitmap(max_db_size, or_rule_indices, batch_size):
    """Host OR-policy bitmap (actor.rs:3601-3625): row q (stride
    ceil(max_db_size / 64) words) has the bits of query q's OR-rule DB ids."""
    stride = -(-max_db_size // 64)
    bm = np.zeros(stride * batch_size, np.uint64)
    for q, ids in enumerate(or_rule_indices):
        for d in ids:
            bm[q * stride + d // 64] |= np.uint64(1) << np.uint64(d % 64)
    return bm


class TwoStageResult:
    """Per party: final_results (first match per query or 0xFFFFFFFF),
    match_counters, all_matches (up to 256 ids per query) and the per-eye
    partial lists; the candidate subset; the matched rotations per eye
    ({query: {db id: [rotations]}})."""

    def __init__(self):
        self.candidates = None
        self.final_results, self.match_counters, self.all_matches = [], [], []
        self.partial = {"left": [], "right": []}
        self.rotations = {"left": [], "right": []}


class TwoStageMatcher:
    """Drives the three parties' engines of both eyes through the two-stage
    flow. `eyes` maps "left"/"right" to three PartyEngines each (their DB
    slices hold that eye's shares); the subset slices are allocated here, once,
    and shared by both eyes."""

    def __init__(self, eyes, max_db_size, max_batch, full_scan_side="left", streams=None):
        import torch

        assert full_scan_side in EYES
        self.eyes, self.max_db_size, self.max_batch = eyes, max_db_size, max_batch
        self.full_scan_side = full_scan_side
        self.streams = streams or [stream_create() for _ in range(3)]
        left = eyes["left"]
        self.code_subs = [left[p].codes.alloc_db(DB_CHUNK_SIZE) for p in range(3)]
        self.mask_subs = [left[p].masks.alloc_db(DB_CHUNK_SIZE) for p in range(3)]
        words = -(-max_db_size * max_batch // 64) + 2
        self.bitmaps = {e: [torch.zeros(words, dtype=torch.int64, device="cuda") for _ in range(3)]
                        for e in EYES}
        bwords = -(-max_batch * max_batch // 64) + 2  # batch rows max_batch bits apart
        self.batch_bitmaps = {e: [torch.zeros(bwords, dtype=torch.int64, device="cuda") for _ in range(3)]
                              for e in EYES}

    def _sync(self):
        for s in self.streams:
            stream_synchronize(s)

    def _clear(self, eye, batch=False):
        import torch

        for b in (self.batch_bitmaps if batch else self.bitmaps)[eye]:
            b.zero_()
        torch.cuda.synchronize()

    def match(self, queries, db_size, batch_size, or_rule_indices=None, reauth_targets=(),
              self_compare=False, anon=None):
        """queries[eye] = (code query sets, mask query sets), three each.
        or_rule_indices: per batch query a list of DB ids (or None);
        reauth_targets: DB ids; self_compare: also deduplicate within the batch
        (query sets uploaded from raw shares); anon: distance capture for the
        anonymized statistics, a dict with "buffers" (eye -> three
        anon_stats.DistanceBuffers), "batch_id" and "targets" (db * max_batch +
        query pairs captured even when not matching). The full-scan eye is
        captured in its 0.375 scan (actor.rs:2476-2530), the other eye in a
        0.375 pass over the subset (actor.rs:2040-2120). Returns a
        TwoStageResult."""
        import torch

        assert batch_size <= self.max_batch and db_size <= self.max_db_size
        side1 = self.full_scan_side
        side2 = "right" if side1 == "left" else "left"
        res = TwoStageResult()
        for e in EYES:
            self._clear(e)
            if self_compare:
                self._clear(e, batch=True)

        def capture(eye, on):
            for p in range(3):
                cmp = self.eyes[eye][p].comparator
                if on:
                    anon["buffers"][eye][p].reset()
                    cmp.set_anon_capture(anon["buffers"][eye][p], anon.get("batch_id", 0),
                                         targets=anon.get("targets", ()))
                else:
                    cmp.set_anon_capture(None)

        # phase 1: full scan of one eye at the prefilter threshold
        qc, qm = queries[side1]
        if anon:
            capture(side1, True)
        scan_db_trio(self.eyes[side1], qc, qm, db_size, batch_size, THRESHOLD_ANON_STATS_A,
                     [b.data_ptr() for b in self.bitmaps[side1]], self.streams)
        self._sync()
        if anon:
            capture(side1, False)

        # candidate subset (actor.rs:938-970)
        or_ids = [] if or_rule_indices is None else [d for ids in or_rule_indices for d in ids]
        extra = sorted(set(int(d) for d in or_ids) | set(int(d) for d in reauth_targets))
        cands = [self.eyes[side1][p].comparator.get_partial_results(
            self.bitmaps[side1][p].data_ptr(), db_size, batch_size, extra, stream=self.streams[p])
            for p in range(3)]
        for p in (1, 2):
            if not np.array_equal(cands[p], cands[0]):
                raise RuntimeError(f"parties 0 and {p} opened different candidate subsets")
        res.candidates = cands[0]
        ids = [self.eyes[side1][p].comparator.partial_ids_ptr() for p in range(3)]
        self._clear(side1)  # actor.rs:971-980: the prefilter bitmap is reused

        # phase 2: both eyes against the subset at the matching threshold
        n = int(res.candidates.size)
        for eye in (side1, side2):
            qc, qm = queries[eye]
            if self_compare:  # actor.rs:1929-1935: batch dedup first
                scan_batch_self_trio(self.eyes[eye], qc, qm, batch_size, THRESHOLD_A,
                                     [b.data_ptr() for b in self.batch_bitmaps[eye]], self.streams,
                                     row_bits=self.max_batch)
            stats = bool(anon) and eye == side2
            if stats:
                capture(eye, True)
            scan_subset_trio(self.eyes[eye], self.code_subs, self.mask_subs, qc, qm, ids, n, db_size,
                             batch_size, THRESHOLD_A, [b.data_ptr() for b in self.bitmaps[eye]],
                             self.streams, anon_threshold=THRESHOLD_ANON_STATS_A if stats else 0)
            self._sync()
            if stats:
                capture(eye, False)
            res.rotations[eye] = [self.eyes[eye][p].comparator.get_partial_results_with_rotations(
                self.streams[p]) for p in range(3)]

        # merge (actor.rs:1097-1131)
        or_dev = None
        if or_rule_indices is not None and any(len(x) for x in or_rule_indices):
            assert len(or_rule_indices) == batch_size
            host = prepare_or_policy_bitmap(self.max_db_size, or_rule_indices, self.max_batch)
            or_dev = torch.from_numpy(host.view(np.int64)).cuda()
            torch.cuda.synchronize()
        for p in range(3):
            cmp = self.eyes["left"][p].comparator
            s = self.streams[p]
            final = torch.full((self.max_batch,), -1, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            cmp.reset(s)
            cmp.join_db_matches(self.bitmaps["left"][p].data_ptr(), self.bitmaps["right"][p].data_ptr(),
                                final.data_ptr(), db_size, or_dev.data_ptr() if or_dev is not None else None,
                                self.max_db_size, s)
            if self_compare:
                cmp.join_batch_matches(self.batch_bitmaps["left"][p].data_ptr(),
                                       self.batch_bitmaps["right"][p].data_ptr(), final.data_ptr(),
                                       batch_size, self.max_batch, s)
            stream_synchronize(s)
            mc, am, mcl, mcr, prl, prr = cmp.buffers()
            B = self.max_batch
            res.final_results.append(final.cpu().numpy().view(np.uint32)[:batch_size].copy())
            res.match_counters.append(download(mc, B, np.uint32, s)[:batch_size])
            res.all_matches.append(download(am, 256 * B, np.uint32, s).reshape(B, 256)[:batch_size])
            res.partial["left"].append((download(mcl, B, np.uint32, s)[:batch_size],
                                        download(prl, 256 * B, np.uint32, s).reshape(B, 256)[:batch_size]))
            res.partial["right"].append((download(mcr, B, np.uint32, s)[:batch_size],
                                         download(prr, 256 * B, np.uint32, s).reshape(B, 256)[:batch_size]))
        return res
