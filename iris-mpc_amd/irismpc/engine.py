"""Python host-side mirror of the reference engines, over the C ABI.

Names and argument meanings follow iris-mpc-gpu:
  ShareDB            src/dot/share_db.rs:448-1070
  Circuits           src/threshold_ring/protocol.rs:433-2611
  DistanceComparator src/dot/distance_comparator.rs:28-894
Device buffers are passed as integer device pointers (e.g. torch tensors'
data_ptr()); streams as integer hipStream_t handles.
"""
import ctypes

import numpy as np

from ._lib import call, ptr_array

IRIS_CODE_LENGTH = 12800
MASK_CODE_LENGTH = 6400
ROTATIONS = 31
DB_CHUNK_SIZE = 1 << 15
# dot/mod.rs:10-15: A = floor((1 - 2 t) * 2^16)
THRESHOLD_A = int((1.0 - 2.0 * 0.345) * (1 << 16))
THRESHOLD_ANON_STATS_A = int((1.0 - 2.0 * 0.375) * (1 << 16))


def _out():
    return ctypes.c_void_p()


def _u32x8(seed):
    a = np.ascontiguousarray(np.asarray(seed, dtype=np.uint32).reshape(8))
    return a, a.ctypes.data


def derive_seed(seed, nonce):
    """server/actor.rs:2914-2925."""
    s, sp = _u32x8(seed)
    out = np.zeros(8, np.uint32)
    call("irismpc_derive_seed", sp, nonce, out.ctypes.data)
    return out


def stream_create(device=0):
    s = _out()
    call("irismpc_stream_create", device, ctypes.byref(s))
    return s.value


def stream_synchronize(stream):
    call("irismpc_stream_synchronize", stream)


def stream_destroy(stream):
    call("irismpc_stream_destroy", stream)


def download(ptr, n, dtype, stream=0):
    """Copies n elements of dtype from a device pointer to a new numpy array."""
    out = np.zeros(n, dtype)
    call("irismpc_memcpy_dtoh", out.ctypes.data, ptr, out.nbytes, stream)
    return out


def fill_u64(ptr, n, value=0, stream=0):
    """n u64 words at ptr := value, by a kernel on `stream` (graph-capture safe)."""
    call("irismpc_fill_u64", ptr, n, value, stream)


def upload(ptr, arr, stream=0):
    a = np.ascontiguousarray(arr)
    call("irismpc_memcpy_htod", ptr, a.ctypes.data, a.nbytes, stream)


class Exchange:
    """Replaces the three parties' NcclComm on one GPU (helpers/comm.rs)."""

    def __init__(self, device=0, timeout_s=300.0):
        h = _out()
        call("irismpc_exchange_create", device, timeout_s, ctypes.byref(h))
        self.h = h.value
        self.device = device

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_exchange_destroy", self.h)
            self.h = None


class SlicedProcessedDatabase:
    """HBM-resident DB share of one matrix (share_db.rs:114-184)."""

    def __init__(self, engine, h):
        self.engine = engine
        self.h = h

    def load_records(self, index, records, stream=0):
        r = np.ascontiguousarray(records, dtype=np.uint16)
        n = r.size // self.engine.code_length
        assert r.size == n * self.engine.code_length
        call("irismpc_db_load_records", self.engine.h, self.h, index, n, r.ctypes.data, stream)

    def load_single_record_from_db(self, index, record, stream=0):
        self.load_records(index, record, stream)

    def load_single_record_from_s3(self, index, limb0, limb1, stream=0):
        l0 = np.ascontiguousarray(limb0, dtype=np.uint8)
        l1 = np.ascontiguousarray(limb1, dtype=np.uint8)
        n = l0.size // self.engine.code_length
        call("irismpc_db_load_records_limbs", self.engine.h, self.h, index, n, l0.ctypes.data,
             l1.ctypes.data, stream)

    def load_full_db(self, records, stream=0):
        self.load_records(0, records, stream)
        return records.size // self.engine.code_length

    def write_at_index(self, query, src_index, dst_index, stream=0):
        call("irismpc_db_write_at_index", self.engine.h, self.h, query.h, src_index, dst_index, stream)

    def fill_random(self, index, n, key, nonce=0, stream=0):
        k, kp = _u32x8(key)
        call("irismpc_db_fill_random", self.engine.h, self.h, index, n, kp, nonce, stream)

    def load_records_device(self, index, n, dev_rows_ptr, stream=0):
        """n u16 rows already on the device -> DB rows [index, index + n)."""
        call("irismpc_db_load_records_device", self.engine.h, self.h, index, n, dev_rows_ptr, stream)

    def replicate_records(self, src, n, dst, stream=0):
        """Rows [src, src + n) -> [dst, dst + n) on the device (16-row multiples)."""
        call("irismpc_db_replicate_records", self.engine.h, self.h, src, n, dst, stream)

    def prefetch_subset(self, ids_ptr, n, subset, stream=0):
        """prefetch_subset (share_db.rs:344-407): rows ids[0..n) (device u32)
        -> rows 0..n of `subset`."""
        call("irismpc_db_prefetch_subset", self.engine.h, self.h, ids_ptr, n, subset.h, stream)

    def read_records(self, index, n, stream=0):
        out = np.zeros((n, self.engine.code_length), np.uint16)
        call("irismpc_db_read_records", self.engine.h, self.h, index, n, out.ctypes.data, stream)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_db_destroy", self.h)
            self.h = None


class DeviceQuery:
    """Query limbs + sums on the device (CudaVec2DSlicerU8 + query_sums)."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_query_destroy", self.h)
            self.h = None


class ShareDB:
    """share_db.rs:448-1070. comm=None builds a non-remote engine."""

    def __init__(self, peer_id, max_db_length, query_length, code_length, chacha_seeds,
                 comm=None, tag=0, device=0):
        so, sop = _u32x8(chacha_seeds[0])
        sp, spp = _u32x8(chacha_seeds[1])
        h = _out()
        call("irismpc_share_db_init", device, peer_id, max_db_length, query_length, code_length,
             sop, spp, comm.h if comm else None, tag, ctypes.byref(h))
        self.h = h.value
        self.peer_id = peer_id
        self.max_db_length = max_db_length
        self.query_length = query_length
        self.code_length = code_length
        self.comm = comm

    @property
    def is_remote(self):
        """Remote engines (with an exchange) mask their results with ChaCha12."""
        return self.comm is not None

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_share_db_destroy", self.h)
            self.h = None

    def alloc_db(self, max_db_length):
        h = _out()
        call("irismpc_share_db_alloc_db", self.h, max_db_length, ctypes.byref(h))
        return SlicedProcessedDatabase(self, h.value)

    def upload_query(self, rows, stream=0):
        """preprocess_query + htod + query_sums: rows = query_length x code_length u16."""
        r = np.ascontiguousarray(rows, dtype=np.uint16)
        n = r.size // self.code_length
        h = _out()
        call("irismpc_query_upload", self.h, r.ctypes.data, n, stream, ctypes.byref(h))
        return DeviceQuery(h.value)

    def upload_query_shares(self, shares, stream=0):
        """Raw query shares (batch x code_length u16) of this party; Lagrange
        preprocessing and the 31 rotations run on the GPU."""
        r = np.ascontiguousarray(shares, dtype=np.uint16).reshape(-1, self.code_length)
        h = _out()
        call("irismpc_query_upload_shares", self.h, r.ctypes.data, r.shape[0], stream, ctypes.byref(h))
        return DeviceQuery(h.value)

    def upload_query_shares_mirrored(self, shares, insert_shares, stream=0):
        """Mirror-orientation queries: rows from the mirrored `shares` (the other
        eye's raw shares), DB form from insert_shares (this eye's)."""
        r = np.ascontiguousarray(shares, dtype=np.uint16).reshape(-1, self.code_length)
        i = np.ascontiguousarray(insert_shares, dtype=np.uint16).reshape(-1, self.code_length)
        assert i.shape == r.shape
        h = _out()
        call("irismpc_query_upload_shares_mirrored", self.h, r.ctypes.data, i.ctypes.data, r.shape[0],
             stream, ctypes.byref(h))
        return DeviceQuery(h.value)

    def upload_query_shares_rows(self, shares, rows, stream=0):
        """Raw query shares as exactly `rows` query rows (rows >= batch * 31; may
        exceed query_length): a packed query set of the batching form."""
        r = np.ascontiguousarray(shares, dtype=np.uint16).reshape(-1, self.code_length)
        h = _out()
        call("irismpc_query_upload_shares_rows", self.h, r.ctypes.data, r.shape[0], rows, stream,
             ctypes.byref(h))
        return DeviceQuery(h.value)

    def refill_query_shares(self, query, shares, rows=0, stream=0):
        """New raw shares into a query set made by upload_query_shares[_rows],
        reusing its device buffers (rows and batch within what it was made with)."""
        r = np.ascontiguousarray(shares, dtype=np.uint16).reshape(-1, self.code_length)
        call("irismpc_query_refill_shares", self.h, query.h, r.ctypes.data, r.shape[0], rows, stream)
        return query

    def refill_query_shares_async(self, query, shares_ptr, batch, rows, stream):
        """refill_query_shares from pinned host memory at shares_ptr (batch x
        code_length u16), ordered on `stream` without synchronising; every
        asynchronous refill of this engine goes on that one stream."""
        call("irismpc_query_refill_shares_async", self.h, query.h, shares_ptr, batch, rows, stream)
        return query

    def dot_reduce_packed(self, query, db, offset, total, chunk_size, multiplier, rows, out_ptrs,
                          ctr_first=None, stream=0):
        """Several batches packed in `query` (rows[i] rows each) in one launch;
        batch i's results at out_ptrs[i] as its own dot_reduce_range lays them out."""
        n = len(rows)
        r = (ctypes.c_size_t * n)(*rows)
        c = (ctypes.c_uint64 * n)(*ctr_first) if ctr_first is not None else None
        call("irismpc_share_db_dot_reduce_packed", self.h, query.h, db.h, offset, total, chunk_size,
             multiplier, r, ptr_array(out_ptrs), n, c, stream)

    def dot_reduce(self, query, db, chunk_size, offset, multiplier=1, stream=0):
        """dot + dot_reduce_and_multiply (fused)."""
        call("irismpc_share_db_dot_reduce", self.h, query.h, db.h, chunk_size, offset, multiplier, stream)

    def dot_reduce_chunks(self, query, db, offset, total, chunk_size, multiplier=1, out_offset=0,
                          stream=0):
        """Consecutive dot_reduce calls over rows [offset, offset + total) in one
        launch; chunk c's results at element out_offset + c * chunk_size * rows
        of the result buffer, with the masks of those calls."""
        call("irismpc_share_db_dot_reduce_chunks", self.h, query.h, db.h, offset, total, chunk_size,
             multiplier, out_offset, stream)

    def dot_reduce_range(self, query, db, offset, total, chunk_size, multiplier, out_ptr, stream=0):
        call("irismpc_share_db_dot_reduce_range", self.h, query.h, db.h, offset, total, chunk_size,
             multiplier, out_ptr, stream)

    def reshare_results(self, chunk_size, stream=0):
        call("irismpc_share_db_reshare_results", self.h, chunk_size, stream)

    @staticmethod
    def reshare_results_trio(engines, chunk_size, streams):
        call("irismpc_share_db_reshare_results_trio", ptr_array([e.h for e in engines]), chunk_size,
             ptr_array(list(streams)))

    def set_peer_copy(self, copy=True):
        """Reshare into a private copy instead of a view of the previous party's
        result buffer."""
        call("irismpc_share_db_set_peer_copy", self.h, 1 if copy else 0)

    def result_chunk_shares(self):
        a, b = _out(), _out()
        call("irismpc_share_db_result_views", self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def fetch_results(self, n, stream=0):
        out = np.zeros(n, np.uint16)
        call("irismpc_share_db_fetch_results", self.h, out.ctypes.data, n, stream)
        return out

    @property
    def counter(self):
        c = ctypes.c_uint64()
        call("irismpc_share_db_get_counter", self.h, ctypes.byref(c))
        return c.value

    @counter.setter
    def counter(self, v):
        call("irismpc_share_db_set_counter", self.h, v)

    def set_variant(self, v):
        call("irismpc_share_db_set_variant", self.h, v)

    def set_timing(self, on=True):
        """Brackets every share GEMM launch with HIP events on its stream."""
        call("irismpc_share_db_set_timing", self.h, 1 if on else 0)

    def launch_times(self):
        """(timed launches, summed duration in ms) since the previous call."""
        n, ms = ctypes.c_size_t(), ctypes.c_double()
        call("irismpc_share_db_launch_times", self.h, ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value


class Graph:
    """HIP graph of the library calls issued on one stream between capture()
    and end(); launch() replays it."""

    def __init__(self, stream):
        self.stream, self.h = stream, None

    def capture(self):
        call("irismpc_graph_capture_begin", self.stream)
        return self

    def end(self):
        h = _out()
        call("irismpc_graph_capture_end", self.stream, ctypes.byref(h))
        self.h = h.value
        return self

    def launch(self, stream=None):
        call("irismpc_graph_launch", self.h, self.stream if stream is None else stream)

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_graph_destroy", self.h)
            self.h = None


def gemm_variants():
    n = ctypes.c_int()
    call("irismpc_gemm_variants", ctypes.byref(n))
    return n.value


def gemm_rot_variants():
    """Ids of the rotation-aware GEMM variants (queries from raw shares only)."""
    first, n = ctypes.c_int(), ctypes.c_int()
    call("irismpc_gemm_rot_variants", ctypes.byref(first), ctypes.byref(n))
    return list(range(first.value, first.value + n.value))


class Circuits:
    """threshold_ring/protocol.rs:433-2611."""

    def __init__(self, peer_id, input_size, alloc_size, chacha_seeds, comm, tag=0, device=0):
        so, sop = _u32x8(chacha_seeds[0])
        sp, spp = _u32x8(chacha_seeds[1])
        h = _out()
        call("irismpc_circuits_new", device, peer_id, input_size, alloc_size, sop, spp, comm.h, tag,
             ctypes.byref(h))
        self.h = h.value
        self.peer_id = peer_id
        self.comm = comm

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_circuits_destroy", self.h)
            self.h = None

    def set_timing(self, on=True):
        """Times each threshold comparison's circuit launches (party 0's engine
        keeps the intervals of the three parties' joint launches)."""
        call("irismpc_circuits_set_timing", self.h, 1 if on else 0)

    def launch_times(self):
        n, ms = ctypes.c_size_t(), ctypes.c_double()
        call("irismpc_circuits_launch_times", self.h, ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    def set_chunk_size(self, words):
        call("irismpc_circuits_set_chunk_size", self.h, words)

    def compare_threshold_masked_many(self, code, mask, length, threshold_a, stream=0):
        """code/mask = (a_ptr, b_ptr) replicated u16 shares. Collective."""
        call("irismpc_circuits_compare_threshold_masked_many", self.h, code[0], code[1], mask[0],
             mask[1], length, threshold_a, stream)

    @staticmethod
    def compare_threshold_masked_many_trio(circuits, codes, masks, length, threshold_a, streams):
        call("irismpc_circuits_compare_threshold_masked_many_trio",
             ptr_array([c.h for c in circuits]), ptr_array([c[0] for c in codes]),
             ptr_array([c[1] for c in codes]), ptr_array([m[0] for m in masks]),
             ptr_array([m[1] for m in masks]), length, threshold_a, ptr_array(list(streams)))

    def take_result_buffer(self):
        a, b = _out(), _out()
        call("irismpc_circuits_result_buffer", self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def open(self, stream=0):
        call("irismpc_circuits_open", self.h, stream)

    @staticmethod
    def open_trio(circuits, streams):
        call("irismpc_circuits_open_trio", ptr_array([c.h for c in circuits]), ptr_array(list(streams)))

    def open_views(self):
        a, b, c = _out(), _out(), _out()
        call("irismpc_circuits_open_views", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    @property
    def counter(self):
        c = ctypes.c_uint64()
        call("irismpc_circuits_get_counter", self.h, ctypes.byref(c))
        return c.value

    @counter.setter
    def counter(self, v):
        call("irismpc_circuits_set_counter", self.h, v)

    def set_lift_capture(self, a_ptr, b_ptr):
        call("irismpc_circuits_set_lift_capture", self.h, a_ptr, b_ptr)


class DistanceComparator:
    """dot/distance_comparator.rs:28-894."""

    def __init__(self, query_length, max_db_size, device=0):
        h = _out()
        call("irismpc_comparator_init", device, query_length, max_db_size, ctypes.byref(h))
        self.h = h.value
        self.query_length = query_length
        self.max_db_size = max_db_size

    def __del__(self):
        if getattr(self, "h", None):
            call("irismpc_comparator_destroy", self.h)
            self.h = None

    def open_results(self, a, b, c, bitmap_ptr, chunk_len, real_chunk_len, offset, total_db_len,
                     batch_size, stream=0):
        call("irismpc_comparator_open_results", self.h, a, b, c, bitmap_ptr, chunk_len,
             real_chunk_len, offset, total_db_len, batch_size, stream)

    def join_db_matches(self, left, right, final_results, db_size, or_policy=None, max_db_size=0,
                        stream=0):
        call("irismpc_comparator_join_db_matches", self.h, left, right, final_results, db_size,
             or_policy, max_db_size or self.max_db_size, stream)

    def reset(self, stream=0):
        call("irismpc_comparator_reset", self.h, stream)

    def get_partial_results(self, bitmap_ptr, db_size, batch_size, extra_ids=(), row_bits=0,
                            stream=0):
        """get_partial_results (distance_comparator.rs:545-648) with the OR-rule /
        reauth indices appended as the actor does (actor.rs:946-970): ascending
        unique candidate DB ids (numpy u32). The list also stays on the device
        (partial_ids_ptr)."""
        extra = np.ascontiguousarray(np.asarray(extra_ids, dtype=np.uint32))
        n = ctypes.c_size_t()
        call("irismpc_comparator_get_partial_results", self.h, bitmap_ptr, db_size, batch_size,
             row_bits, extra.ctypes.data if extra.size else None, extra.size, None, 0,
             ctypes.byref(n), stream)
        out = np.zeros(n.value, np.uint32)
        if n.value:
            call("irismpc_comparator_get_partial_results", self.h, bitmap_ptr, db_size, batch_size,
                 row_bits, extra.ctypes.data if extra.size else None, extra.size, out.ctypes.data,
                 out.size, ctypes.byref(n), stream)
        return out

    def join_batch_matches(self, left, right, final_results, batch_size, row_bits=0, stream=0):
        """join_batch_matches (distance_comparator.rs:731-746) over batch
        self-comparison bitmaps (bit row_bits * q + j)."""
        call("irismpc_comparator_join_batch_matches", self.h, left, right, final_results, batch_size,
             row_bits, stream)

    def set_anon_capture(self, buffers=None, batch_id=0, max_query_length=0, max_db_length=0,
                         targets=()):
        """Match-distance capture of the following opens into `buffers` (an
        anon_stats.DistanceBuffers); None switches it off."""
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.uint64))
        if buffers is None:
            call("irismpc_comparator_set_anon_capture", self.h, None, None, None, None, None, None, 0,
                 0, 0, 0, None, 0)
            return
        call("irismpc_comparator_set_anon_capture", self.h, *buffers.pointers(), buffers.capacity,
             batch_id, max_query_length or self.query_length, max_db_length or self.max_db_size,
             t.ctypes.data if t.size else None, t.size)

    def partial_ids_ptr(self):
        p = _out()
        call("irismpc_comparator_partial_ids", self.h, ctypes.byref(p))
        return p.value

    def open_results_with_index_mapping(self, a, b, c, bitmap_ptr, chunk_len, real_chunk_len,
                                        total_db_len, batch_size, index_map_ptr, stream=0):
        call("irismpc_comparator_open_results_with_index_mapping", self.h, a, b, c, bitmap_ptr,
             chunk_len, real_chunk_len, total_db_len, batch_size, index_map_ptr, stream)

    def get_partial_results_with_rotations(self, stream=0):
        """distance_comparator.rs:650-711: {query: {db id: [rotations]}} of the
        index-mapped opens since the last call (the capture is reset)."""
        cap = 256 * self.query_length
        q = np.zeros(cap, np.uint32)
        d = np.zeros(cap, np.uint32)
        r = np.zeros(cap, np.int8)
        n = ctypes.c_size_t()
        call("irismpc_comparator_fetch_partial_rotations", self.h, q.ctypes.data, d.ctypes.data,
             r.ctypes.data, cap, ctypes.byref(n), stream)
        out = {}
        for qi, di, ri in zip(q[:min(n.value, cap)].tolist(), d[:min(n.value, cap)].tolist(),
                              r[:min(n.value, cap)].tolist()):
            out.setdefault(qi, {}).setdefault(di, []).append(ri)
        return out

    def buffers(self):
        outs = [_out() for _ in range(6)]
        call("irismpc_comparator_buffers", self.h, *[ctypes.byref(o) for o in outs])
        return [o.value for o in outs]


def partial_results_noalloc(bitmap_ptr, db_size, batch_size, ids_ptr, count_ptr, extra_ptr=None, n_extra=0,
                            row_bits=0, stream=0):
    """The candidate list of DistanceComparator.get_partial_results on the
    device without allocating or synchronising (irismpc_partial_results_noalloc):
    ids_ptr (device u32, room for db_size ids) and count_ptr (device u32)."""
    call("irismpc_partial_results_noalloc", bitmap_ptr, db_size, batch_size, row_bits, extra_ptr, n_extra,
         ids_ptr, count_ptr, stream)


class PartyEngines:
    """One party's engines for a DB scan (the ServerActor's codes/masks engines,
    phase-2 Circuits and DistanceComparator, with HBM-resident DB shares)."""

    def __init__(self, codes, masks, circuits, comparator, code_db, mask_db):
        self.codes, self.masks, self.circuits, self.comparator = codes, masks, circuits, comparator
        self.code_db, self.mask_db = code_db, mask_db


def scan_db(party, q_code, q_mask, db_size, batch_size, threshold_a, bitmap_ptr, stream=0):
    """compare_query_against_db_and_self_prefiltering's chunk loop for one party
    (server/actor.rs:2246-2589). Collective: call from the three party threads."""
    call("irismpc_scan_db", party.codes.h, party.masks.h, party.circuits.h, party.comparator.h,
         party.code_db.h, party.mask_db.h, q_code.h, q_mask.h, db_size, batch_size, threshold_a,
         bitmap_ptr, stream)


def scan_chunk_trio(parties, q_codes, q_masks, offset, db_size, batch_size, threshold_a, bitmap_ptrs,
                    streams, slot_ptrs=None, row_bits=0):
    """One chunk step of the three parties' scans (rows [offset, offset + 32768)
    of a db_size-row DB); slot_ptrs[p] (device u32 per batch query) picks each
    query's bitmap row, rows row_bits bits apart (0: db_size)."""
    call("irismpc_scan_chunk_trio", ptr_array([p.codes.h for p in parties]),
         ptr_array([p.masks.h for p in parties]), ptr_array([p.circuits.h for p in parties]),
         ptr_array([p.comparator.h for p in parties]), ptr_array([p.code_db.h for p in parties]),
         ptr_array([p.mask_db.h for p in parties]), ptr_array([q.h for q in q_codes]),
         ptr_array([q.h for q in q_masks]), offset, db_size, batch_size, threshold_a,
         ptr_array(list(bitmap_ptrs)), ptr_array(list(slot_ptrs)) if slot_ptrs is not None else None,
         row_bits, ptr_array(list(streams)))


def scan_chunks_trio(parties, q_codes, q_masks, offset, n_chunks, db_size, batch_size, threshold_a,
                     bitmap_ptrs, streams, slot_ptrs=None, row_bits=0):
    """n_chunks consecutive chunk steps from `offset` as one step (the dot
    shares of all of them fit the engines' result buffers); the opened bits
    are those of n_chunks calls of scan_chunk_trio."""
    call("irismpc_scan_chunks_trio", ptr_array([p.codes.h for p in parties]),
         ptr_array([p.masks.h for p in parties]), ptr_array([p.circuits.h for p in parties]),
         ptr_array([p.comparator.h for p in parties]), ptr_array([p.code_db.h for p in parties]),
         ptr_array([p.mask_db.h for p in parties]), ptr_array([q.h for q in q_codes]),
         ptr_array([q.h for q in q_masks]), offset, n_chunks, db_size, batch_size, threshold_a,
         ptr_array(list(bitmap_ptrs)), ptr_array(list(slot_ptrs)) if slot_ptrs is not None else None,
         row_bits, ptr_array(list(streams)))


def scan_subset_trio(parties, code_subs, mask_subs, q_codes, q_masks, ids_ptrs, n_ids, total_db_len,
                     batch_size, threshold_a, bitmap_ptrs, streams, anon_threshold=0):
    """The DB half of compare_query_against_db_subset_and_self (actor.rs:
    1900-2244) for the three parties: subset ids (device u32 per party) are
    gathered into the subset slices, compared at threshold_a and opened at
    their DB ids."""
    call("irismpc_scan_subset_trio", ptr_array([p.codes.h for p in parties]),
         ptr_array([p.masks.h for p in parties]), ptr_array([p.circuits.h for p in parties]),
         ptr_array([p.comparator.h for p in parties]), ptr_array([p.code_db.h for p in parties]),
         ptr_array([p.mask_db.h for p in parties]), ptr_array([d.h for d in code_subs]),
         ptr_array([d.h for d in mask_subs]), ptr_array([q.h for q in q_codes]),
         ptr_array([q.h for q in q_masks]), ptr_array(list(ids_ptrs)), n_ids, total_db_len,
         batch_size, threshold_a, anon_threshold, ptr_array(list(bitmap_ptrs)), ptr_array(list(streams)))


def scan_batch_self_trio(parties, q_codes, q_masks, batch_size, threshold_a, bitmap_ptrs, streams,
                         row_bits=0):
    """compare_query_against_self (actor.rs:1789-1897) for the three parties:
    bit row_bits * q + j = query q matches query j (query sets uploaded from
    raw shares)."""
    call("irismpc_scan_batch_self_trio", ptr_array([p.codes.h for p in parties]),
         ptr_array([p.masks.h for p in parties]), ptr_array([p.circuits.h for p in parties]),
         ptr_array([p.comparator.h for p in parties]), ptr_array([q.h for q in q_codes]),
         ptr_array([q.h for q in q_masks]), batch_size, threshold_a, ptr_array(list(bitmap_ptrs)),
         row_bits, ptr_array(list(streams)))


def scan_db_trio(parties, q_codes, q_masks, db_size, batch_size, threshold_a, bitmap_ptrs, streams):
    call("irismpc_scan_db_trio", ptr_array([p.codes.h for p in parties]),
         ptr_array([p.masks.h for p in parties]), ptr_array([p.circuits.h for p in parties]),
         ptr_array([p.comparator.h for p in parties]), ptr_array([p.code_db.h for p in parties]),
         ptr_array([p.mask_db.h for p in parties]), ptr_array([q.h for q in q_codes]),
         ptr_array([q.h for q in q_masks]), db_size, batch_size, threshold_a,
         ptr_array(list(bitmap_ptrs)), ptr_array(list(streams)))
