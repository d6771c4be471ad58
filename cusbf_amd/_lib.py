"""ctypes binding of libcusbf_hip.so (the C ABI declared in include/cusbf_hip.h).

The library is built in-tree by cusbf_amd.build. Loading fails loudly when it
is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcusbf_hip.so")

P = ctypes.c_void_p
u8 = ctypes.c_uint8
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int
sz = ctypes.c_size_t


class CConfig(ctypes.Structure):
    _fields_ = [
        ("k", u32),
        ("s", u32),
        ("m", u32),
        ("hash_count", u32),
        ("alphabet", u32),
        ("symbol_count", u32),
        ("invalid_symbol", u8),
        ("separator", u8),
        ("reserved", u8 * 2),
        ("lut", u8 * 256),
    ]


class CRecordRange(ctypes.Structure):
    _fields_ = [("sequence_offset", u64), ("sequence_bytes", u64)]


class CInsertReport(ctypes.Structure):
    _fields_ = [("records_indexed", u64), ("indexed_bases", u64), ("inserted_kmers", u64)]


class CServerOptions(ctypes.Structure):
    _fields_ = [("staging_bytes", u64), ("slots", u32), ("max_wait_us", u32)]


class CServerStats(ctypes.Structure):
    _fields_ = [("requests", u64), ("batches", u64), ("max_requests_per_batch", u64),
                ("bytes", u64), ("windows", u64), ("fill_seconds", ctypes.c_double),
                ("launch_seconds", ctypes.c_double), ("device_seconds", ctypes.c_double),
                ("copy_seconds", ctypes.c_double)]


class CQueryReport(ctypes.Structure):
    _fields_ = [("records_queried", u64), ("queried_bases", u64), ("queried_kmers", u64),
                ("positive_kmers", u64)]


class CBatchTuning(ctypes.Structure):
    _fields_ = [("waves_per_simd", u32), ("blocks_per_cu", u32), ("claim", u32),
                ("reserved", u32)]


class CFastxRecordView(ctypes.Structure):
    _fields_ = [("record_index", u64), ("header", ctypes.c_void_p), ("header_len", u64),
                ("sequence", ctypes.c_void_p), ("sequence_len", u64), ("queried_bases", u64),
                ("queried_kmers", u64), ("positive_kmers", u64), ("hits", ctypes.c_void_p),
                ("hits_len", u64)]


FASTX_CONSUMER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(CFastxRecordView), ctypes.c_void_p)
READ_FN = ctypes.CFUNCTYPE(ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64)

# (name, restype, argtypes) of every exported entry point of include/cusbf_hip.h.
CFG = ctypes.POINTER(CConfig)
RR = ctypes.POINTER(CRecordRange)
SIGNATURES = [
    ("cusbf_last_error", i32, [ctypes.c_char_p, sz]),
    ("cusbf_abi_version", i32, []),
    ("cusbf_config_init", i32, [CFG, u32, u32, u32, u32, u32]),
    ("cusbf_config_symbol_width", i32, [CFG, ctypes.POINTER(u32)]),
    ("cusbf_config_symbol_bits", i32, [CFG, ctypes.POINTER(u32)]),
    ("cusbf_create", i32, [CFG, u64, ctypes.POINTER(P)]),
    ("cusbf_destroy", None, [P]),
    ("cusbf_clear", i32, [P, P]),
    ("cusbf_load_factor", i32, [P, ctypes.POINTER(ctypes.c_float)]),
    ("cusbf_filter_bits", u64, [P]),
    ("cusbf_num_shards", u64, [P]),
    ("cusbf_words_device_ptr", i32, [P, ctypes.POINTER(P)]),
    ("cusbf_copy_words_to_host", i32, [P, P, u64]),
    ("cusbf_insert_sequence_async", i32, [P, P, u64, P, ctypes.POINTER(u64)]),
    ("cusbf_insert_sequence", i32, [P, P, u64, P, ctypes.POINTER(u64)]),
    ("cusbf_contains_sequence_async", i32, [P, P, u64, P, u64, P]),
    ("cusbf_contains_sequence", i32, [P, P, u64, P, u64, P]),
    ("cusbf_contains_sequence_host_to_device", i32, [P, P, u64, P, u64, P]),
    ("cusbf_insert_dense_packed", i32, [P, P, u64, u64, P, ctypes.POINTER(u64)]),
    ("cusbf_contains_dense_packed_host_to_device", i32, [P, P, u64, u64, P, u64, P]),
    ("cusbf_contains_dense_packed", i32, [P, P, u64, u64, P, u64, P]),
    ("cusbf_dense_packed_word_count", u64, [CFG, u64]),
    ("cusbf_dense_packed_kmer_count", u64, [CFG, u64]),
    ("cusbf_pack_dense_sequence_async", i32, [CFG, P, u64, P, u64, P]),
    ("cusbf_insert_dense_packed_async", i32, [P, P, u64, u64, P, ctypes.POINTER(u64)]),
    ("cusbf_contains_dense_packed_async", i32, [P, P, u64, u64, P, u64, P]),
    ("cusbf_insert_record_batch_async", i32, [P, P, u64, RR, u64, P, ctypes.POINTER(u64)]),
    ("cusbf_contains_record_batch_async", i32, [P, P, u64, RR, u64, P, u64, P]),
    ("cusbf_insert_record_batch", i32, [P, P, u64, RR, u64, P, ctypes.POINTER(CInsertReport)]),
    ("cusbf_query_record_batch", i32, [P, P, u64, RR, u64, P, ctypes.POINTER(CQueryReport)]),
    ("cusbf_query_record_batch_hits", i32,
     [P, P, u64, RR, u64, P, P, u64, ctypes.POINTER(CQueryReport)]),
    ("cusbf_count_positive_async", i32, [P, u64, P, P]),
    ("cusbf_count_positive_per_record_async", i32, [P, P, P, u64, P, P]),
    ("cusbf_count_windows_runs_async", i32, [P, P, u64, P, P]),
    ("cusbf_kernel_family", ctypes.c_char_p, [P]),
    ("cusbf_set_insert_path", i32, [P, u32, u32, u32]),
    ("cusbf_get_insert_path", i32, [P, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                    ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    ("cusbf_set_sorted_workspace_limit", i32, [P, u64]),
    ("cusbf_get_sorted_workspace", i32, [P, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                         ctypes.POINTER(u32)]),
    ("cusbf_batch_variants", u32, [P]),
    ("cusbf_batch_bucket", u32, [u64, u64]),
    ("cusbf_batch_tuning_set", i32, [P, u32, u32, ctypes.POINTER(CBatchTuning)]),
    ("cusbf_batch_tuning_get", i32, [P, u32, u32, ctypes.POINTER(CBatchTuning)]),
    ("cusbf_insert_batched_async", i32, [P, P, u64, P, u64, P]),
    ("cusbf_contains_batched_async", i32, [P, P, u64, P, u64, P, u64, P]),
    ("cusbf_batched_status", i32, [P, P]),
    ("cusbf_join_records_async", i32, [P, P, u64, P, u64, P, u64, P]),
    ("cusbf_contains_batched_noalloc_async", i32, [P, P, u64, P, u64, P, u64, P]),
    ("cusbf_batch_plan_create", i32, [P, u32, P, u64, P, u64, P, u64,
                                      ctypes.POINTER(CBatchTuning), ctypes.POINTER(P)]),
    ("cusbf_batch_plan_launch", i32, [P, P]),
    ("cusbf_batch_plan_status", i32, [P, P]),
    ("cusbf_batch_plan_destroy", None, [P]),
    ("cusbf_count_batched_windows_runs_async", i32, [P, P, u64, P, u64, P, P]),
    ("cusbf_insert_fastx_buffer", i32, [P, P, u64, ctypes.c_char_p, ctypes.c_double, P,
                                        ctypes.POINTER(CInsertReport)]),
    ("cusbf_insert_fastx_reader", i32, [P, READ_FN, P, ctypes.c_char_p, ctypes.c_double, P,
                                        ctypes.POINTER(CInsertReport)]),
    ("cusbf_insert_fastx_file", i32, [P, ctypes.c_char_p, ctypes.c_double, P,
                                      ctypes.POINTER(CInsertReport)]),
    ("cusbf_query_fastx_buffer", i32, [P, P, u64, ctypes.c_char_p, ctypes.c_double, P,
                                       FASTX_CONSUMER, P, ctypes.POINTER(CQueryReport)]),
    ("cusbf_query_fastx_reader", i32, [P, READ_FN, P, ctypes.c_char_p, ctypes.c_double, P,
                                       FASTX_CONSUMER, P, ctypes.POINTER(CQueryReport)]),
    ("cusbf_query_fastx_file", i32, [P, ctypes.c_char_p, ctypes.c_double, P, FASTX_CONSUMER, P,
                                     ctypes.POINTER(CQueryReport)]),
    ("cusbf_parse_fastx_buffer", i32, [P, u64, ctypes.c_char_p, FASTX_CONSUMER, P,
                                       ctypes.POINTER(u64)]),
    ("cusbf_parse_fastx_buffer_chunked", i32, [P, u64, ctypes.c_char_p, u64, u64, FASTX_CONSUMER, P,
                                               ctypes.POINTER(u64)]),
    ("cusbf_fastx_set_chunk_bytes", i32, [P, u64]),
    ("cusbf_last_error_location", i32, [ctypes.c_char_p, sz, ctypes.POINTER(u32),
                                        ctypes.POINTER(u32)]),
    ("cusbf_save", i32, [P, ctypes.c_char_p]),
    ("cusbf_image_config", i32, [ctypes.c_char_p, CFG, ctypes.POINTER(u64)]),
    ("cusbf_load", i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    ("cusbf_server_create", i32, [P, ctypes.POINTER(CServerOptions), ctypes.POINTER(P)]),
    ("cusbf_server_destroy", None, [P]),
    ("cusbf_server_submit", i32, [P, P, u64, RR, u64, P, u64, ctypes.POINTER(u64)]),
    ("cusbf_server_request_windows", u64, [P, RR, u64]),
    ("cusbf_server_wait", i32, [P, u64, ctypes.POINTER(CQueryReport),
                                ctypes.POINTER(ctypes.c_double)]),
    ("cusbf_server_stats_get", i32, [P, ctypes.POINTER(CServerStats)]),
]

_lib = None


def share_hip_runtime() -> None:
    """Load torch's HIP runtime before any library of this package.

    torch ships its own libamdhip64 with the soname libamdhip64.so.7. Loaded
    first, it satisfies this package's libraries by soname, so the process
    holds one HIP runtime and device pointers and streams are shared. Loading
    these libraries first would bring in a second runtime next to torch's.
    """
    try:
        import torch  # noqa: F401
    except ImportError:  # C/C++ callers without torch use the system runtime
        pass


def load(path: str = LIB_PATH):
    """Load libcusbf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first (python -m cusbf_amd.build)")
    share_hip_runtime()
    lib = ctypes.CDLL(path)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib
