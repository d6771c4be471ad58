"""cusbf_amd: MI355X-native Super Bloom filter (HIP kernels behind a C ABI).

Python mirror of the reference's cusbf::filter<Config> hot-path API; the
compute runs in libcusbf_hip.so (built in-tree by cusbf_amd.build).
"""
from .config import (  # noqa: F401
    Alphabet,
    Config,
    DnaAlphabet,
    DnaTripletAlphabet,
    ProteinAlphabet,
)
from .errors import CusbfError, ErrorCategory, Result  # noqa: F401
from .filter import (  # noqa: F401
    INSERT_ATOMIC,
    INSERT_AUTO,
    INSERT_SORTED,
    Filter,
    InsertPathInfo,
    SortedWorkspaceInfo,
    InsertReport,
    QueryReport,
    RecordRange,
    count_positive,
    encode_symbols,
    pack_dense_sequence,
    pack_dense_sequence_device,
)
from .normalize import NormalizedRecord, normalize_record_batch, query_layout  # noqa: F401
from .serve import Server, query_record_batch_hits  # noqa: F401

__all__ = [
    "Alphabet", "Config", "DnaAlphabet", "DnaTripletAlphabet", "ProteinAlphabet", "CusbfError",
    "ErrorCategory", "Result", "Filter", "InsertReport", "QueryReport", "RecordRange",
    "count_positive", "encode_symbols", "pack_dense_sequence", "pack_dense_sequence_device",
    "NormalizedRecord", "normalize_record_batch", "query_layout", "Server",
    "query_record_batch_hits", "INSERT_AUTO", "INSERT_ATOMIC", "INSERT_SORTED", "InsertPathInfo",
    "SortedWorkspaceInfo",
]
