"""Batch processing of the reference GPU server, three parties on one GPU:
ServerActor::process_batch_query (iris-mpc-gpu/src/server/actor.rs:674-1790)
in the Normal orientation, over the HBM-resident DB of both eyes.

Per batch:
- drop invalid entries (actor.rs:829-850);
- deletions overwrite their rows with the dummy iris, then identity updates
  write their new shares over their rows (apply_non_mpc_updates,
  actor.rs:2632-2700); a batch of updates only stops there;
- LUC look-back: the last luc_lookback_records DB rows join every request's
  OR-rule ids except reauths and reset / recovery checks (generate_luc_records,
  actor.rs:772-796,3623-3659);
- two-stage matching with batch deduplication (the two_stage module);
- result assembly: merged results, match ids (DB ids, and 0xFFFFFFFF - j for
  an earlier query j of the batch), partial per-eye matches with their
  rotations, uniqueness insertions (no match, under the supermatch threshold,
  not skip_persistence; reset and recovery checks are matched but never
  inserted), successful reauths (actor.rs:1300-1480;
  get_merged_results, distribute_insertions, calculate_insertion_indices
  :3174-3303);
- DB writes: insertions at the end of the DB, then successful reauths at their
  targets, both eyes (write_db_at_index, actor.rs:1522-1600,3547-3599).

process_job runs the full-face mirror-attack pass first when the batch asks
for it (actor.rs:585-640): the Mirror orientation queries each eye's DB with
the other eye's shares mirrored, writes nothing, and its matches both flag
attacks (no Normal match but a Mirror match) and veto insertions in the
Normal pass that follows.

Lists that the reference fills in atomic-arrival order on the device (match
ids, partial ids, rotations) are returned sorted.
"""
from dataclasses import dataclass, field

import numpy as np

from .engine import (DB_CHUNK_SIZE, IRIS_CODE_LENGTH, MASK_CODE_LENGTH, Circuits,
                     DistanceComparator, Exchange, PartyEngines, ShareDB, derive_seed)
from .two_stage import EYES, TwoStageMatcher

NON_MATCH_ID = 0xFFFFFFFF                 # actor.rs:192
SUPERMATCH_THRESHOLD = 4000               # actor.rs:117
UNIQUENESS_MESSAGE_TYPE = "uniqueness"    # iris-mpc-common/src/helpers/smpc_request.rs:109
REAUTH_MESSAGE_TYPE = "reauth"            # smpc_request.rs:110
RESET_CHECK_MESSAGE_TYPE = "reset_check"  # smpc_request.rs:111
RECOVERY_CHECK_MESSAGE_TYPE = "recovery_check"  # smpc_request.rs:112
NORMAL, MIRROR = "normal", "mirror"       # Orientation (actor.rs:119-134)
ALL_MATCHES_LEN = 256


def get_merged_results(host_results, n_devices):
    """actor.rs:3174-3193: per query the smallest global id over devices
    (device-local id * n_devices + device), NON_MATCH_ID if none."""
    out = []
    for j in range(len(host_results[0])):
        best = NON_MATCH_ID
        for i in range(len(host_results)):
            r = int(host_results[i][j])
            idx = (r * n_devices + i) & 0xFFFFFFFF
            if r != NON_MATCH_ID and idx < best:
                best = idx
        out.append(best)
    return out


def distribute_insertions(results, db_sizes):
    """actor.rs:3194-3206: round-robin over devices, from the smallest DB."""
    ret = [[] for _ in db_sizes]
    c = db_sizes.index(min(db_sizes))
    for r in results:
        ret[c].append(r)
        c = (c + 1) % len(db_sizes)
    return ret


def calculate_insertion_indices(merged_results, insertion_list, db_sizes, batch_size):
    """actor.rs:3254-3288: inserted queries get consecutive new global ids
    (from the total DB size, visiting devices round-robin from the one whose
    first insertion is earliest) in merged_results; returns `matches` (False
    exactly for the inserted queries)."""
    matches = [True] * batch_size
    last = sum(db_sizes)
    min_index, min_val = 0, None
    for i, lst in enumerate(insertion_list):
        if lst and (min_val is None or lst[0] < min_val):
            min_val, min_index = lst[0], i
    c = 0
    while True:
        for i in range(len(insertion_list)):
            idx = (i + min_index) % len(insertion_list)
            if c >= len(insertion_list[idx]):
                return matches
            q = insertion_list[idx][c]
            merged_results[q] = last
            matches[q] = False
            last += 1
        c += 1


def generate_luc_records(latest_luc_index, or_rule_indices, lookback_records, batch_size,
                         skip_lookback_requests=()):
    """actor.rs:3623-3659: the ids latest - lookback .. latest (saturating at 0)
    join each request's OR-rule ids (then sorted, without repeats), except the
    requests listed in skip_lookback_requests; with no OR-rule ids at all every
    request gets the look-back ids alone; lookback_records = 0 changes
    nothing."""
    if lookback_records == 0:
        return [list(x) for x in or_rule_indices]
    start = max(0, latest_luc_index - lookback_records)
    ids = list(range(start, latest_luc_index + 1))
    if not or_rule_indices:
        return [list(ids) for _ in range(batch_size)]
    out = []
    for idx, (or_ids, luc) in enumerate(zip(or_rule_indices, [ids] * batch_size)):
        if idx in skip_lookback_requests:
            out.append(list(or_ids))
        else:
            out.append(sorted(set(or_ids) | set(luc)))
    return out


@dataclass
class BatchQuery:
    """The fields of BatchQuery (iris-mpc-common/src/job.rs:34-93) that the
    matching path reads. left/right: per party (code shares [B][12800],
    mask shares [B][6400]) u16, raw (as stored in the DB)."""
    request_ids: list
    request_types: list
    left: list
    right: list
    or_rule_indices: list = field(default_factory=list)
    skip_persistence: list = None
    reauth_target_indices: dict = field(default_factory=dict)
    deletion_requests_indices: list = field(default_factory=list)
    valid_entries: list = None
    full_face_mirror_attacks_detection_enabled: bool = False
    luc_lookback_records: int = 0
    # identity updates (job.rs:103-107,152-170): new shares of whole DB rows,
    # per update (left, right), each per party (code [12800], mask [6400])
    identity_update_request_ids: list = field(default_factory=list)
    identity_update_indices: list = field(default_factory=list)
    identity_update_shares: list = field(default_factory=list)

    def retain(self, keep):
        """BatchQuery::retain: only the entries at positions `keep`."""
        k = list(keep)
        sel = lambda sh: [(c[k], m[k]) for c, m in sh]  # noqa: E731
        ids = [self.request_ids[i] for i in k]
        return BatchQuery(
            ids, [self.request_types[i] for i in k], sel(self.left), sel(self.right),
            [self.or_rule_indices[i] for i in k] if self.or_rule_indices else [],
            [self.skip_persistence[i] for i in k] if self.skip_persistence is not None else None,
            {r: t for r, t in self.reauth_target_indices.items() if r in ids},
            list(self.deletion_requests_indices), None, self.full_face_mirror_attacks_detection_enabled,
            self.luc_lookback_records, list(self.identity_update_request_ids),
            list(self.identity_update_indices), list(self.identity_update_shares))


@dataclass
class ServerJobResult:
    """ServerJobResult (iris-mpc-common/src/job.rs:365-435), the matching
    fields."""
    merged_results: list
    request_ids: list
    request_types: list
    matches: list
    matches_with_skip_persistence: list
    skip_persistence: list
    match_ids: list
    partial_match_ids_left: list
    partial_match_ids_right: list
    partial_match_rotation_indices_left: list
    partial_match_rotation_indices_right: list
    partial_match_counters_left: list
    partial_match_counters_right: list
    deleted_ids: list
    matched_batch_request_ids: list
    successful_reauths: list
    reauth_target_indices: dict
    candidates: np.ndarray = None
    full_face_mirror_attack_detected: list = field(default_factory=list)
    full_face_mirror_match_ids: list = field(default_factory=list)
    full_face_mirror_partial_match_ids_left: list = field(default_factory=list)
    full_face_mirror_partial_match_ids_right: list = field(default_factory=list)
    full_face_mirror_partial_match_counters_left: list = field(default_factory=list)
    full_face_mirror_partial_match_counters_right: list = field(default_factory=list)
    one_sided_match_distances: dict = None   # eye -> per party OneSidedDistanceCache
    two_sided_match_distances: list = None   # per party TwoSidedDistanceCache
    identity_update_request_ids: list = field(default_factory=list)
    identity_update_indices: list = field(default_factory=list)


def _dummy_shares():
    """Per-party shares of the deletion dummy: the default iris, code all
    zeros and mask all ones (get_dummy_shares_for_deletion,
    iris-mpc-common/src/iris_db/mod.rs:10-23); the share randomness comes
    from this library's generator instead of StdRng(0), the secret is the
    same."""
    from . import synth

    code = np.zeros((1, 200), np.uint64)
    mask = np.full((1, 200), np.uint64(0xFFFFFFFFFFFFFFFF))
    sh = synth.shares_of(code, mask, 0x0de1e7e, 0)
    return [(sh[p][0], sh[p][1]) for p in range(3)]


class ServerActor:
    """Three parties' engines for both eyes, the HBM DB of both eyes and the
    batch loop. seeds: per party the (own, prev) 8-word ChaCha seeds (default:
    party p = all words p); engine seeds are derived from them as the
    reference's actor derives its engines' seeds (derive_seed nonces 0, 1, 5
    for codes, masks, circuits)."""

    def __init__(self, max_db_size, max_batch_size, full_scan_side="left", seeds=None,
                 return_partial_results=True, full_scan_side_switching_enabled=False,
                 match_distances_buffer_size=0):
        self.max_db_size, self.max_batch_size = max_db_size, max_batch_size
        # anonymized statistics: capture buffers per orientation, eye and party
        # (DistanceCache, anon_stats.rs:13-57); 0 = no capture
        self.match_distances_buffer_size = match_distances_buffer_size
        self.internal_batch_counter = 0
        self.anon_buffers = None
        if match_distances_buffer_size:
            from .anon_stats import DistanceBuffers

            self.anon_buffers = {o: {e: [DistanceBuffers(match_distances_buffer_size) for _ in range(3)]
                                     for e in EYES} for o in (NORMAL, MIRROR)}
        self.full_scan_side_switching_enabled = full_scan_side_switching_enabled
        self.return_partial_results = return_partial_results
        Qr = max_batch_size * 31
        if seeds is None:
            base = [np.full(8, p, np.uint32) for p in range(3)]
            seeds = [(base[p], base[(p + 2) % 3]) for p in range(3)]
        self.exchanges, eyes = [], {}
        for k, eye in enumerate(EYES):
            ex = Exchange(timeout_s=300)
            self.exchanges.append(ex)
            parties = []
            for p in range(3):
                so, sp = seeds[p]
                der = lambda n: (derive_seed(so, n + 16 * k), derive_seed(sp, n + 16 * k))  # noqa: E731
                codes = ShareDB(p, DB_CHUNK_SIZE, Qr, IRIS_CODE_LENGTH, der(0), comm=ex, tag=0)
                masks = ShareDB(p, DB_CHUNK_SIZE, Qr, MASK_CODE_LENGTH, der(1), comm=ex, tag=1)
                circ = Circuits(p, Qr * DB_CHUNK_SIZE, Qr * DB_CHUNK_SIZE // 64, der(5), ex, tag=2)
                cmp = DistanceComparator(Qr, max_db_size)
                parties.append(PartyEngines(codes, masks, circ, cmp, codes.alloc_db(max_db_size),
                                            masks.alloc_db(max_db_size)))
            eyes[eye] = parties
        self.eyes = eyes
        self.matcher = TwoStageMatcher(eyes, max_db_size, max_batch_size, full_scan_side)
        self.current_db_size = 0
        self._dummy = None

    @property
    def full_scan_side(self):
        return self.matcher.full_scan_side

    def load_db(self, left, right, index=0):
        """DB rows [index, index + n) of both eyes from per-party raw shares
        ((code [n][12800], mask [n][6400]) per party)."""
        n = left[0][0].shape[0]
        assert index + n <= self.max_db_size
        for eye, sh in (("left", left), ("right", right)):
            for p in range(3):
                self.eyes[eye][p].code_db.load_records(index, sh[p][0])
                self.eyes[eye][p].mask_db.load_records(index, sh[p][1])
        self.current_db_size = max(self.current_db_size, index + n)

    def load_stored_records(self, chunks, max_serial_id=0):
        """Party p's DB-chunk bytes chunks[p] (the store module's record
        format) straight into HBM; the DB size grows to the largest loaded
        serial id (loader.rs:196-240)."""
        from .store import load_stored_records

        ends = []
        for p in range(3):
            L, R = self.eyes["left"][p], self.eyes["right"][p]
            n, end = load_stored_records(L.codes, L.masks, L.code_db, L.mask_db, R.code_db, R.mask_db,
                                         chunks[p], max_serial_id)
            ends.append(end)
        self.current_db_size = max([self.current_db_size] + ends)
        return ends

    def _write_db_at_index(self, queries, src_index, dst_index):
        for eye in EYES:
            qc, qm = queries[eye]
            for p in range(3):
                pe = self.eyes[eye][p]
                pe.code_db.write_at_index(qc[p], src_index, dst_index)
                pe.mask_db.write_at_index(qm[p], src_index, dst_index)

    def _apply_deletions(self, indices):
        if not indices:
            return
        if self._dummy is None:
            self._dummy = _dummy_shares()
        for d in indices:
            if d >= self.current_db_size:
                continue  # actor.rs:2650-2656: out of bounds, skipped
            for eye in EYES:
                for p in range(3):
                    self.eyes[eye][p].code_db.load_records(d, self._dummy[p][0])
                    self.eyes[eye][p].mask_db.load_records(d, self._dummy[p][1])

    def _apply_identity_updates(self, batch):
        """actor.rs:2681-2720: each update's new shares overwrite its DB row in
        both eyes; rows beyond the DB are skipped."""
        for idx, (left, right) in zip(batch.identity_update_indices, batch.identity_update_shares):
            if idx >= self.current_db_size:
                continue
            for eye, sh in (("left", left), ("right", right)):
                for p in range(3):
                    self.eyes[eye][p].code_db.load_records(idx, np.asarray(sh[p][0]).reshape(1, -1))
                    self.eyes[eye][p].mask_db.load_records(idx, np.asarray(sh[p][1]).reshape(1, -1))

    def _distances(self, batch, orientation, skip, targets):
        """One-sided caches per eye and party (load_additions_since,
        actor.rs:2196-2240,2555-2589) and their two-sided merge
        (actor.rs:1631-1650)."""
        from .anon_stats import RECOVERY, REAUTH, UNIQUENESS, TwoSidedDistanceCache, load_additions

        ops = [REAUTH if k == REAUTH_MESSAGE_TYPE else RECOVERY if k == "recovery_check" else UNIQUENESS
               for k in batch.request_types]
        qmax = self.max_batch_size * 31
        one = {e: [load_additions(self.anon_buffers[orientation][e][p], 0, self.match_distances_buffer_size,
                                  ops, skip, targets, qmax, self.max_db_size) for p in range(3)]
               for e in EYES}
        two = [TwoSidedDistanceCache.merge(one["left"][p], one["right"][p]) for p in range(3)]
        return one, two

    def process_job(self, batch):
        """The actor's job loop body (actor.rs:585-640): with mirror-attack
        detection, the Mirror orientation first, then the Normal one with the
        mirror results; the full-scan side alternates when switching is on."""
        if batch.full_face_mirror_attacks_detection_enabled:
            mirror = self.process_batch_query(batch, MIRROR, None)
            res = self.process_batch_query(batch, NORMAL, mirror)
        else:
            res = self.process_batch_query(batch, NORMAL, None)
        if self.full_scan_side_switching_enabled:
            self.matcher.full_scan_side = "right" if self.matcher.full_scan_side == "left" else "left"
        return res

    def _queries(self, batch, orientation):
        """Device query sets per eye. Mirror (server/mod.rs:156-176): eye E's
        DB is queried with the other eye's shares mirrored; the batch
        self-comparison still uses E's own shares as its DB side."""
        queries = {}
        for eye, other in (("left", "right"), ("right", "left")):
            own = getattr(batch, eye)
            qc, qm = [], []
            for p in range(3):
                codes, masks = self.eyes[eye][p].codes, self.eyes[eye][p].masks
                if orientation == MIRROR:
                    mir = getattr(batch, other)
                    qc.append(codes.upload_query_shares_mirrored(mir[p][0], own[p][0]))
                    qm.append(masks.upload_query_shares_mirrored(mir[p][1], own[p][1]))
                else:
                    qc.append(codes.upload_query_shares(own[p][0]))
                    qm.append(masks.upload_query_shares(own[p][1]))
            queries[eye] = (qc, qm)
        return queries

    def process_batch_query(self, batch, orientation=NORMAL, previous_results=None):
        assert orientation in (NORMAL, MIRROR)
        if batch.valid_entries is not None and not all(batch.valid_entries):
            batch = batch.retain([i for i, v in enumerate(batch.valid_entries) if v])
        B = len(batch.request_ids)
        assert B <= self.max_batch_size
        skip = list(batch.skip_persistence) if batch.skip_persistence is not None else [False] * B
        if previous_results is None:  # actor.rs:2637-2639: only the first pass of a job
            self._apply_deletions(list(batch.deletion_requests_indices))
            self._apply_identity_updates(batch)
        if B == 0:  # non-MPC updates only (actor.rs:709-720)
            return ServerJobResult([], [], [], [], [], [], [], [], [], [], [], [], [],
                                   list(batch.deletion_requests_indices), [], [], {},
                                   identity_update_request_ids=list(batch.identity_update_request_ids),
                                   identity_update_indices=list(batch.identity_update_indices))

        queries = self._queries(batch, orientation)
        or_rule = [list(x) for x in batch.or_rule_indices] if batch.or_rule_indices else []
        if orientation == NORMAL and (or_rule or batch.luc_lookback_records > 0):
            skip_luc = {i for i, k in enumerate(batch.request_types)
                        if k in (REAUTH_MESSAGE_TYPE, RESET_CHECK_MESSAGE_TYPE, RECOVERY_CHECK_MESSAGE_TYPE)}
            or_rule = generate_luc_records(max(self.current_db_size, 1) - 1, or_rule,
                                           batch.luc_lookback_records, B, skip_luc)
        or_rule = or_rule or None
        self.internal_batch_counter += 1  # actor.rs:860
        anon = None
        if self.anon_buffers is not None:
            # reauth pairs are captured even when not matching (actor.rs:838-858)
            targets = [i + batch.reauth_target_indices[rid] * self.max_batch_size
                       for i, (kind, rid) in enumerate(zip(batch.request_types, batch.request_ids))
                       if kind == REAUTH_MESSAGE_TYPE and rid in batch.reauth_target_indices]
            anon = {"buffers": self.anon_buffers[orientation], "batch_id": self.internal_batch_counter,
                    "targets": targets}
        r = self.matcher.match(queries, self.current_db_size, B, or_rule_indices=or_rule,
                               reauth_targets=list(batch.reauth_target_indices.values()),
                               self_compare=True, anon=anon)
        for p in (1, 2):
            if not (np.array_equal(r.final_results[p], r.final_results[0])
                    and np.array_equal(r.match_counters[p], r.match_counters[0])):
                raise RuntimeError(f"parties 0 and {p} disagree on the merged results")

        merged = get_merged_results([r.final_results[0].tolist()], 1)
        counters = r.match_counters[0]
        match_ids = [sorted(r.all_matches[0][j][:min(int(counters[j]), ALL_MATCHES_LEN)].tolist())
                     for j in range(B)]
        partial_ids, partial_cnt, rot_idx = {}, {}, {}
        for eye in EYES:
            cnt, lists = r.partial[eye][0]
            partial_cnt[eye] = [int(c) for c in cnt]
            partial_ids[eye] = [sorted(lists[j][:min(int(cnt[j]), ALL_MATCHES_LEN)].tolist()) for j in range(B)]
            rots = r.rotations[eye][0]
            rot_idx[eye] = [[sorted(rots.get(j, {}).get(d, [])) for d in partial_ids[eye][j]
                             if d in rots.get(j, {})] for j in range(B)]

        # mirror-attack detection (actor.rs:1304-1374)
        prev = previous_results if orientation == NORMAL else None
        if prev is not None:
            attack = [merged[i] == NON_MATCH_ID and prev.matches_with_skip_persistence[i]
                      and batch.request_types[i] == UNIQUENESS_MESSAGE_TYPE for i in range(B)]
        else:
            attack = [False] * B

        # uniqueness insertions (actor.rs:1383-1417); with mirror results, only
        # entries that did not match in the mirror orientation either
        uniq = [i for i in range(B)
                if batch.request_types[i] == UNIQUENESS_MESSAGE_TYPE and merged[i] == NON_MATCH_ID
                and partial_cnt["left"][i] <= SUPERMATCH_THRESHOLD
                and partial_cnt["right"][i] <= SUPERMATCH_THRESHOLD
                and not (prev is not None and prev.matches_with_skip_persistence[i])]
        inserts = [i for i in uniq if not skip[i]]
        skipped = [i for i in uniq if skip[i]]
        insertion_list = distribute_insertions(inserts, [self.current_db_size])
        matches = calculate_insertion_indices(merged, insertion_list, [self.current_db_size], B)
        mwsp = list(matches)
        for i in skipped:
            mwsp[i] = False
        matched_batch = [[batch.request_ids[NON_MATCH_ID - x] for x in ids if x > NON_MATCH_ID - B]
                         for ids in match_ids]
        if previous_results is not None:  # actor.rs:1431-1450: union with the mirror pass's
            matched_batch = [ids + [x for x in m if x not in ids]
                             for ids, m in zip(matched_batch, previous_results.matched_batch_request_ids)]
        filtered = [[x for x in ids if x <= NON_MATCH_ID - self.max_batch_size] for ids in match_ids]
        reauth_ok = [batch.request_types[i] == REAUTH_MESSAGE_TYPE
                     and batch.reauth_target_indices[batch.request_ids[i]] in filtered[i] for i in range(B)]

        # DB writes, Normal orientation only (actor.rs:1480-1600)
        if orientation == NORMAL:
            n_ins = sum(len(x) for x in insertion_list)
            if self.current_db_size + n_ins > self.max_db_size:
                raise RuntimeError("DB size exceeded")
            for i in insertion_list[0]:
                self._write_db_at_index(queries, i, self.current_db_size)
                self.current_db_size += 1
            for i in range(B):
                if reauth_ok[i] and not skip[i]:
                    t = batch.reauth_target_indices[batch.request_ids[i]]
                    if t < self.current_db_size:
                        self._write_db_at_index(queries, i, t)

        def per_eye(d):
            return d["left"], d["right"]

        pil, pir = per_eye(partial_ids)
        ril, rir = per_eye(rot_idx) if self.return_partial_results else ([[]] * B, [[]] * B)
        pcl, pcr = per_eye(partial_cnt)
        # the result's match ids leave out the batch matches (actor.rs:1451-1458,1686)
        res = ServerJobResult(merged, list(batch.request_ids), list(batch.request_types), matches,
                              mwsp, skip, filtered, pil, pir, ril, rir, pcl, pcr,
                              list(batch.deletion_requests_indices), matched_batch, reauth_ok,
                              dict(batch.reauth_target_indices), r.candidates,
                              identity_update_request_ids=list(batch.identity_update_request_ids),
                              identity_update_indices=list(batch.identity_update_indices))
        res.full_face_mirror_attack_detected = attack
        if anon is not None:
            res.one_sided_match_distances, res.two_sided_match_distances = self._distances(
                batch, orientation, skip, anon["targets"])
        if previous_results is not None:
            res.full_face_mirror_match_ids = previous_results.match_ids
            res.full_face_mirror_partial_match_ids_left = previous_results.partial_match_ids_left
            res.full_face_mirror_partial_match_ids_right = previous_results.partial_match_ids_right
            res.full_face_mirror_partial_match_counters_left = previous_results.partial_match_counters_left
            res.full_face_mirror_partial_match_counters_right = previous_results.partial_match_counters_right
        return res
