"""Record-batch normalisation and query layout (host logic).

normalize_record_batch follows include/cusbf/normalized_record_batch.hpp:77-304:
records are validated (ordered, non-overlapping, in range, symbol aligned),
joined with the alphabet separator (padded to the symbol width), and each
record's valid k-mer count is recorded. query_layout follows
include/cusbf/detail/query_layout.cuh:41-72: per-record hit offsets in the
normalised buffer and the total hit count including boundary windows.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, List, Sequence, Tuple

import numpy as np

from .config import Config
from .errors import CusbfError, ErrorCategory
from .filter import RecordRange, _host_bytes, encode_symbols


@dataclass
class NormalizedRecord:
    """normalized_record_batch.hpp:20-29"""

    record_index: int
    input_offset: int
    size: int
    valid_kmers: int


@dataclass
class QueryLayoutRecord:
    """query_layout.cuh:13-20"""

    record_index: int
    input_offset: int
    size: int
    valid_kmers: int
    hit_offset: int
    hit_count: int


def _kmer_count(config: Config, nbytes: int) -> int:
    s = nbytes // config.symbol_width
    return 0 if s < config.k else s - config.k + 1


def valid_record_kmer_count(config: Config, record: np.ndarray) -> int:
    """normalized_record_batch.hpp:87-110 (windows without an invalid symbol)."""
    w = config.symbol_width
    nsym = record.size // w
    if nsym < config.k:
        return 0
    codes = encode_symbols(config, record[: nsym * w])
    bad = (codes == config.alphabet.invalid_symbol).astype(np.int64)
    csum = np.concatenate([[0], np.cumsum(bad)])
    windows = nsym - config.k + 1
    inv = csum[config.k:config.k + windows] - csum[:windows]
    return int(np.count_nonzero(inv == 0))


def validate_record_batch(config: Config, nbytes: int, ranges: Sequence[RecordRange],
                          exceeds_message: str = "record batch range exceeds the source sequence"):
    """normalized_record_batch.hpp:123-147"""
    next_offset = 0
    for r in ranges:
        if r.sequence_offset < next_offset:
            raise CusbfError(ErrorCategory.invalid_argument,
                             "record batch ranges must be ordered and non-overlapping")
        if r.sequence_offset > nbytes or r.sequence_bytes > nbytes - r.sequence_offset:
            raise CusbfError(ErrorCategory.invalid_argument, exceeds_message)
        w = config.symbol_width
        if r.sequence_offset % w or r.sequence_bytes % w:
            raise CusbfError(ErrorCategory.invalid_argument,
                             "record batch ranges must align to the configured alphabet symbol width")
        next_offset = r.sequence_offset + r.sequence_bytes


def normalize_record_batch(config: Config, sequence, ranges: Iterable
                           ) -> Tuple[bytes, List[NormalizedRecord]]:
    """normalized_record_batch.hpp:288-316: (normalised bytes, per-record metadata)."""
    b = _host_bytes(sequence)
    rr = [r if isinstance(r, RecordRange) else RecordRange(int(r[0]), int(r[1])) for r in ranges]
    validate_record_batch(config, b.size, rr)
    w = config.symbol_width
    sep = bytes([config.alphabet.separator])
    parts: List[bytes] = []
    size = 0
    records: List[NormalizedRecord] = []
    for i, r in enumerate(rr):
        if size != 0:  # write_record_boundary (:208-232)
            rem = size % w
            if rem:
                parts.append(sep * (w - rem))
                size += w - rem
            parts.append(sep * w)
            size += w
        rec = b[r.sequence_offset:r.sequence_offset + r.sequence_bytes]
        parts.append(rec.tobytes())
        size += rec.size
        records.append(NormalizedRecord(i, r.sequence_offset, int(rec.size),
                                        valid_record_kmer_count(config, rec)))
    return b"".join(parts), records


def query_layout(config: Config, records: Sequence[NormalizedRecord]
                 ) -> Tuple[List[QueryLayoutRecord], int]:
    """query_layout.cuh:41-72: (per-record hit spans, total hit count)."""
    w = config.symbol_width
    logical = 0
    out: List[QueryLayoutRecord] = []
    for r in records:
        if logical != 0:
            rem = logical % w
            if rem:
                logical += w - rem
            logical += w
        out.append(QueryLayoutRecord(r.record_index, r.input_offset, r.size, r.valid_kmers,
                                     logical // w, _kmer_count(config, r.size)))
        logical += r.size
    return out, _kmer_count(config, logical)
