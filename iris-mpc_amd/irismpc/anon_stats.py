"""Anonymized-statistics distance caches (iris-mpc-gpu/src/server/anon_stats.rs).

The opens of the matching scans can capture, per opened match at the
statistics threshold (0.375) and per requested reauth pair, the party's four
dot-product shares (code a/b, mask a/b) with a match id
    batch_id * max_db * max_query + db * max_query + query row.
DistanceBuffers are one eye's device buffers of one party; load_additions
turns them into a OneSidedDistanceCache keyed by match id / 31 (one entry per
(batch, db, query), the rotations in its list); TwoSidedDistanceCache keeps
the keys captured in both eyes. Bucket statistics over these distances are
computed by a separate service and are not part of this library.
"""
from dataclasses import dataclass

import numpy as np

ROTATIONS = 31
UNIQUENESS, REAUTH, RECOVERY = "uniqueness", "reauth", "recovery"  # AnonStatsOperation


@dataclass(frozen=True)
class CpuDistanceShare:
    """anon_stats.rs:297-305."""
    idx: int
    code_a: int
    code_b: int
    mask_a: int
    mask_b: int
    operation: str


class DistanceBuffers:
    """One eye's capture buffers of one party (DistanceCache::init,
    anon_stats.rs:29-57): `capacity` entries on the device."""

    def __init__(self, capacity):
        import torch

        self.capacity = capacity
        mk = lambda dt: torch.empty(capacity + 1, dtype=dt, device="cuda")  # noqa: E731
        self.code_a, self.code_b, self.mask_a, self.mask_b = (mk(torch.int16) for _ in range(4))
        self.idx = mk(torch.int64)
        self.counter = torch.zeros(2, dtype=torch.int32, device="cuda")
        self.reset()

    def pointers(self):
        return (self.code_a.data_ptr(), self.code_b.data_ptr(), self.mask_a.data_ptr(),
                self.mask_b.data_ptr(), self.idx.data_ptr(), self.counter.data_ptr())

    def reset(self):
        """reset_anon_stats_distance_cache (actor.rs:3304-3318): counter 0,
        indices and shares all ones."""
        import torch

        self.counter.zero_()
        self.idx.fill_(-1)
        for t in (self.code_a, self.code_b, self.mask_a, self.mask_b):
            t.fill_(-1)
        torch.cuda.synchronize()

    def fetch(self):
        """(entries seen, idx u64, code_a, code_b, mask_a, mask_b u16) on the host."""
        import torch

        torch.cuda.synchronize()
        n = int(self.counter[0].item())
        m = min(n, self.capacity)
        host = lambda t, dt: t[:m].cpu().numpy().view(dt)  # noqa: E731
        return (n, host(self.idx, np.uint64), host(self.code_a, np.uint16), host(self.code_b, np.uint16),
                host(self.mask_a, np.uint16), host(self.mask_b, np.uint16))


def operation_from_match_id(idx, operations, max_query_length):
    """anon_stats.rs:307-321."""
    if not operations or max_query_length == 0:
        return UNIQUENESS
    request = (idx % max_query_length) // ROTATIONS
    return operations[request] if request < len(operations) else UNIQUENESS


def target_idx_from_match_id(idx, max_query_length, max_db_size):
    """anon_stats.rs:323-326: db * max_batch + query of a match id."""
    return (idx % (max_query_length * max_db_size)) // ROTATIONS


class OneSidedDistanceCache:
    """anon_stats.rs:328-349: match id / 31 -> the captured shares."""

    def __init__(self, entries=None):
        self.map = entries or {}

    def __len__(self):
        return len(self.map)


class TwoSidedDistanceCache:
    """anon_stats.rs:351-403."""

    def __init__(self, entries=None):
        self.map = entries or {}

    @staticmethod
    def merge(left, right):
        """Keys captured in both eyes, with both eyes' shares."""
        return TwoSidedDistanceCache({k: (v, right.map[k]) for k, v in left.map.items() if k in right.map})

    def extend(self, other):
        for k, v in other.map.items():
            assert k not in self.map, f"Key {k} is not allowed to already exist in the distance cache"
            self.map[k] = v

    def __len__(self):
        return len(self.map)


def load_additions(buffers, old_counter, max_internal_buffer_size, operations, skip_persistence,
                   reauth_targets, max_query_length, max_db_size):
    """DistanceCache::load_additions_since (anon_stats.rs:136-295) for one
    party and eye: entries sorted by match id, those after old_counter kept,
    skip_persistence requests dropped, reauth entries kept only for their
    target pair; grouped by match id / 31. A buffer filled past
    max_internal_buffer_size yields nothing."""
    n, idx, ca, cb, ma, mb = buffers.fetch()
    if n >= max_internal_buffer_size:
        return OneSidedDistanceCache()
    order = np.argsort(idx, kind="stable")
    out = {}
    targets = set(int(t) for t in reauth_targets)
    for k in order[old_counter:n]:
        i = int(idx[k])
        op = operation_from_match_id(i, operations, max_query_length)
        if max_query_length > 0:
            request = (i % max_query_length) // ROTATIONS
            if request < len(skip_persistence) and skip_persistence[request]:
                continue
        if op == REAUTH and target_idx_from_match_id(i, max_query_length, max_db_size) not in targets:
            continue
        out.setdefault(i // ROTATIONS, []).append(
            CpuDistanceShare(i, int(ca[k]), int(cb[k]), int(ma[k]), int(mb[k]), op))
    return OneSidedDistanceCache(out)
