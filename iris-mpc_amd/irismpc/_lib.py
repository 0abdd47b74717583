"""ctypes binding of libirismpc.so (the C ABI declared in include/irismpc.h).

There is no fallback: if the native library is missing or cannot be loaded,
every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libirismpc.so")


class IrisMpcError(RuntimeError):
    pass


_lib = None

vp = ctypes.c_void_p
sz = ctypes.c_size_t
u16 = ctypes.c_uint16
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int
i64 = ctypes.c_int64
dbl = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all return int status)
SIGNATURES = {
    "irismpc_version": [],
    "irismpc_device_count": [vp],
    "irismpc_stream_create": [i32, vp],
    "irismpc_stream_destroy": [vp],
    "irismpc_stream_synchronize": [vp],
    "irismpc_malloc": [i32, sz, vp],
    "irismpc_free": [vp],
    "irismpc_memcpy_htod": [vp, vp, sz, vp],
    "irismpc_memcpy_dtoh": [vp, vp, sz, vp],
    "irismpc_fill_u32": [vp, sz, u32, vp],
    "irismpc_fill_u64": [vp, sz, u64, vp],
    "irismpc_derive_seed": [vp, u64, vp],
    "irismpc_exchange_create": [i32, dbl, vp],
    "irismpc_exchange_destroy": [vp],
    "irismpc_share_db_init": [i32, i32, sz, sz, sz, vp, vp, vp, i32, vp],
    "irismpc_share_db_destroy": [vp],
    "irismpc_share_db_alloc_db": [vp, sz, vp],
    "irismpc_db_destroy": [vp],
    "irismpc_db_load_records": [vp, vp, sz, sz, vp, vp],
    "irismpc_db_load_records_limbs": [vp, vp, sz, sz, vp, vp, vp],
    "irismpc_db_write_at_index": [vp, vp, vp, sz, sz, vp],
    "irismpc_db_read_records": [vp, vp, sz, sz, vp, vp],
    "irismpc_db_fill_random": [vp, vp, sz, sz, vp, u64, vp],
    "irismpc_db_load_records_device": [vp, vp, sz, sz, vp, vp],
    "irismpc_db_replicate_records": [vp, vp, sz, sz, sz, vp],
    "irismpc_query_upload_shares_rows": [vp, vp, sz, sz, vp, vp],
    "irismpc_query_refill_shares": [vp, vp, vp, sz, sz, vp],
    "irismpc_query_refill_shares_async": [vp, vp, vp, sz, sz, vp],
    "irismpc_query_upload_shares_mirrored": [vp, vp, vp, sz, vp, vp],
    "irismpc_share_db_dot_reduce_packed": [vp, vp, vp, sz, sz, sz, u16, vp, vp, i32, vp, vp],
    "irismpc_graph_capture_begin": [vp],
    "irismpc_graph_capture_end": [vp, vp],
    "irismpc_graph_launch": [vp, vp],
    "irismpc_graph_destroy": [vp],
    "irismpc_synth_iris": [u64, u64, vp, vp],
    "irismpc_synth_flip": [vp, u64, u64, u32],
    "irismpc_synth_plain_device": [u64, u64, sz, vp, vp, vp],
    "irismpc_synth_shares_device": [vp, vp, sz, u64, u64, vp, vp, vp],
    "irismpc_query_upload": [vp, vp, sz, vp, vp],
    "irismpc_query_upload_shares": [vp, vp, sz, vp, vp],
    "irismpc_query_destroy": [vp],
    "irismpc_scan_db": [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, u32, vp, vp],
    "irismpc_scan_db_trio": [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, u32, vp, vp],
    "irismpc_scan_chunk_trio": [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, sz, u32, vp, vp, sz, vp],
    "irismpc_scan_chunks_trio": [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, sz, sz, u32, vp, vp, sz, vp],
    "irismpc_share_db_dot_reduce": [vp, vp, vp, sz, sz, u16, vp],
    "irismpc_share_db_dot_reduce_range": [vp, vp, vp, sz, sz, sz, u16, vp, vp],
    "irismpc_share_db_dot_reduce_chunks": [vp, vp, vp, sz, sz, sz, u16, sz, vp],
    "irismpc_share_db_reshare_results": [vp, sz, vp],
    "irismpc_share_db_reshare_results_trio": [vp, sz, vp],
    "irismpc_share_db_result_views": [vp, vp, vp],
    "irismpc_share_db_set_peer_copy": [vp, i32],
    "irismpc_share_db_fetch_results": [vp, vp, sz, vp],
    "irismpc_share_db_get_counter": [vp, vp],
    "irismpc_share_db_set_counter": [vp, u64],
    "irismpc_share_db_set_variant": [vp, i32],
    "irismpc_partial_results_noalloc": [vp, sz, sz, sz, vp, sz, vp, vp, vp],
    "irismpc_share_db_set_timing": [vp, i32],
    "irismpc_share_db_launch_times": [vp, vp, vp],
    "irismpc_circuits_set_timing": [vp, i32],
    "irismpc_circuits_launch_times": [vp, vp, vp],
    "irismpc_gemm_variants": [vp],
    "irismpc_gemm_rot_variants": [vp, vp],
    "irismpc_circuits_new": [i32, i32, sz, sz, vp, vp, vp, i32, vp],
    "irismpc_circuits_destroy": [vp],
    "irismpc_circuits_set_chunk_size": [vp, sz],
    "irismpc_circuits_compare_threshold_masked_many": [vp, vp, vp, vp, vp, sz, u32, vp],
    "irismpc_circuits_compare_threshold_masked_many_trio": [vp, vp, vp, vp, vp, sz, u32, vp],
    "irismpc_circuits_result_buffer": [vp, vp, vp],
    "irismpc_circuits_open": [vp, vp],
    "irismpc_circuits_open_trio": [vp, vp],
    "irismpc_circuits_open_views": [vp, vp, vp, vp],
    "irismpc_circuits_get_counter": [vp, vp],
    "irismpc_circuits_set_counter": [vp, u64],
    "irismpc_circuits_set_lift_capture": [vp, vp, vp],
    "irismpc_comparator_init": [i32, sz, sz, vp],
    "irismpc_comparator_destroy": [vp],
    "irismpc_comparator_open_results": [vp, vp, vp, vp, vp, sz, sz, sz, sz, sz, vp],
    "irismpc_comparator_join_db_matches": [vp, vp, vp, vp, sz, vp, sz, vp],
    "irismpc_comparator_reset": [vp, vp],
    "irismpc_comparator_get_partial_results": [vp, vp, sz, sz, sz, vp, sz, vp, sz, vp, vp],
    "irismpc_comparator_partial_ids": [vp, vp],
    "irismpc_comparator_join_batch_matches": [vp, vp, vp, vp, sz, sz, vp],
    "irismpc_scan_batch_self_trio": [vp, vp, vp, vp, vp, vp, sz, u32, vp, sz, vp],
    "irismpc_comparator_open_results_with_index_mapping": [vp, vp, vp, vp, vp, sz, sz, sz, sz, vp, vp],
    "irismpc_comparator_fetch_partial_rotations": [vp, vp, vp, vp, sz, vp, vp],
    "irismpc_db_prefetch_subset": [vp, vp, vp, sz, vp, vp],
    "irismpc_db_load_stored_records": [vp, vp, vp, vp, vp, vp, vp, sz, u32, vp, vp, vp],
    "irismpc_scan_subset_trio": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, sz, u32, u32, vp, vp],
    "irismpc_comparator_set_anon_capture": [vp, vp, vp, vp, vp, vp, vp, sz, u64, sz, sz, vp, sz],
    "irismpc_comparator_buffers": [vp, vp, vp, vp, vp, vp, vp],
    "irismpc_batch_plan_packs": [vp, sz, sz, sz, vp, vp],
    "irismpc_batch_counter_advance": [sz, sz, sz, vp],
    "irismpc_batch_scan_counter_advance": [sz, sz, sz, vp],
    "irismpc_batch_runner_create": [vp, vp, sz, u16, vp, sz, vp, vp, sz, sz, sz, vp, vp],
    "irismpc_batch_runner_destroy": [vp],
    "irismpc_batch_runner_replan": [vp, sz, vp],
    "irismpc_batch_runner_plan": [vp, vp, vp],
    "irismpc_batch_runner_issue_packed": [vp, u64, sz, vp],
    "irismpc_batch_runner_issue_plain": [vp, sz, vp],
    "irismpc_batch_runner_prepare_persistent": [vp, u64, vp],
    "irismpc_batch_runner_issue_persistent": [vp, vp],
    "irismpc_serving_create": [i32, vp, vp, vp, sz, sz, sz, u32, vp],
    "irismpc_serving_destroy": [vp],
    "irismpc_serving_start": [vp, i32, sz, sz],
    "irismpc_serving_finish": [vp, dbl],
    "irismpc_serving_submit": [vp, i32, vp, vp, i64, dbl, vp],
    "irismpc_serving_result": [vp, u64, vp, vp, vp, sz, vp, vp, vp, vp, sz],
    "irismpc_serving_result_version": [vp, u64, vp],
    "irismpc_serving_result_batch_matches": [vp, u64, vp, sz, vp],
    "irismpc_serving_release": [vp, u64],
    "irismpc_serving_now": [vp, vp],
    "irismpc_serving_stats": [vp, vp],
    "irismpc_serving_loop_times": [vp, vp],
    "irismpc_serving_n_rows": [vp, vp],
    "irismpc_serving_run_alone": [vp, vp, vp, vp, sz, vp],
    "irismpc_serving_time_step": [vp, sz, sz, i32, vp],
    "irismpc_serving_set_admission": [vp, sz, sz],
    "irismpc_serving_set_step_table": [vp, vp, vp, sz],
    "irismpc_serving_tune_steps": [vp, vp, sz, vp, sz, dbl, dbl, vp, vp],
    "irismpc_serving_chunks_per_step": [vp, sz, sz, sz, vp],
    "irismpc_serving_pick_step": [vp, vp, sz, dbl, dbl, vp],
}


def _init_torch_runtime():
    """PyTorch-ROCm carries its own HIP runtime next to the one this library
    links; it has to come up first in the process (the other order leaves the
    library's runtime without a device), so bring it up before loading."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    global _lib
    if _lib is None:
        _init_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise IrisMpcError(
                f"native library {LIB_PATH} is missing: run __graft_entry__.build() "
                "(make -C iris-mpc_amd); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        L.irismpc_last_error.restype = ctypes.c_char_p
        L.irismpc_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = i32
        _lib = L
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().irismpc_last_error().decode(errors="replace")
        raise IrisMpcError(f"{name}: {msg}")
    return rc


def ptr_array(ptrs):
    """C array of 3 (or n) void pointers."""
    return (ctypes.c_void_p * len(ptrs))(*[p if isinstance(p, int) else (p.value if p else None) for p in ptrs])
