"""irismpc: MI355X-native secret-shared iris matching (host side over the C ABI)."""
from ._lib import IrisMpcError, LIB_PATH, lib  # noqa: F401
from .engine import (  # noqa: F401
    DB_CHUNK_SIZE, IRIS_CODE_LENGTH, MASK_CODE_LENGTH, ROTATIONS, THRESHOLD_A,
    THRESHOLD_ANON_STATS_A, Circuits, Graph, PartyEngines, scan_chunk_trio, scan_chunks_trio, scan_db, scan_db_trio, DistanceComparator, Exchange, ShareDB, derive_seed,
    download, fill_u64, gemm_rot_variants, gemm_variants, partial_results_noalloc, stream_create, stream_destroy,
    stream_synchronize,
    upload,
)
