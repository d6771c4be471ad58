"""FASTA / FASTQ ingest: the reference's insert_fastx / query_fastx family.

Reference: include/cusbf/filter.cuh:402-453 (insert), 620-800 (query, per-record
consumer, detailed reports), include/cusbf/Fastx.hpp (records, reports, parser
rules). The library parses the text natively (mmap'd files, gzip files, host
buffers or file-like readers), stages records back to back and runs the batched
kernels chunk by chunk; parse errors raise CusbfError(fastx_parse) whose
`location` is the SourceLocation (file, line, column) and whose message is
"file:line:column: detail".
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Callable, List, Optional

from . import _lib
from .errors import check
from .filter import Filter, InsertReport, QueryReport, stream_handle


@dataclass
class FastxRecordView:
    """Fastx.hpp:172-190 (hits valid only during the consumer call in C++;
    copied here)."""

    record_index: int
    header: str
    sequence: bytes
    queriedBases: int
    queriedKmers: int
    positive_kmers: int
    hits: bytes


@dataclass
class FastxDetailedQueryRecord:
    """Fastx.hpp:193-208."""

    record_index: int
    header: str
    sequence: bytes
    queriedBases: int
    queriedKmers: int
    positive_kmers: int
    hits: List[int] = field(default_factory=list)


@dataclass
class FastxDetailedQueryReport:
    """Fastx.hpp:211-216."""

    summary: QueryReport
    records: List[FastxDetailedQueryRecord]


def _view(rec) -> FastxRecordView:
    r = rec.contents
    header = ctypes.string_at(r.header, r.header_len).decode("latin-1") if r.header_len else ""
    seq = ctypes.string_at(r.sequence, r.sequence_len) if r.sequence_len else b""
    hits = ctypes.string_at(r.hits, r.hits_len) if r.hits_len else b""
    return FastxRecordView(int(r.record_index), header, seq, int(r.queried_bases),
                           int(r.queried_kmers), int(r.positive_kmers), hits)


def _consumer(fn: Optional[Callable[[FastxRecordView], None]]):
    """ctypes consumer wrapping a Python callable; exceptions are kept and re-raised."""
    if fn is None:
        return _lib.FASTX_CONSUMER(), None
    box = {"error": None}

    def call(rec, _user):
        try:
            fn(_view(rec))
            return 0
        except BaseException as e:  # noqa: BLE001 - re-raised after the C call
            box["error"] = e
            return 1  # stops the query

    return _lib.FASTX_CONSUMER(call), box


def _reader(obj):
    """ctypes read callback over a binary file-like object."""
    box = {"error": None}

    def read(_user, dst, cap):
        try:
            data = obj.read(int(cap))
            if isinstance(data, str):
                data = data.encode("latin-1")
            if not data:
                return 0
            ctypes.memmove(dst, data, len(data))
            return len(data)
        except BaseException as e:  # noqa: BLE001
            box["error"] = e
            return -1

    return _lib.READ_FN(read), box


def _text(data):
    if isinstance(data, str):
        data = data.encode("latin-1")
    return bytes(data)


def _report_ins(r) -> InsertReport:
    return InsertReport(int(r.records_indexed), int(r.indexed_bases), int(r.inserted_kmers))


def _report_q(r) -> QueryReport:
    return QueryReport(int(r.records_queried), int(r.queried_bases), int(r.queried_kmers),
                       int(r.positive_kmers))


def _raise_kept(box):
    if box is not None and box.get("error") is not None:
        raise box["error"]


def set_chunk_bytes(filt: Filter, nbytes: int) -> None:
    """Fixed FASTX chunk capacity (0 = from fill_fraction)."""
    lib = _lib.load()
    check(lib, lib.cusbf_fastx_set_chunk_bytes(filt.handle, int(nbytes)))


def insert_fastx(filt: Filter, source, fill_fraction: float = 0.7, stream=None) -> InsertReport:
    """filter.cuh:416-422: `source` is FASTA/FASTQ text (bytes/str) or a binary
    file-like object read to its end."""
    lib = _lib.load()
    rep = _lib.CInsertReport()
    if hasattr(source, "read"):
        fn, box = _reader(source)
        st = lib.cusbf_insert_fastx_reader(filt.handle, fn, None, b"<stream>", fill_fraction,
                                           stream_handle(stream), ctypes.byref(rep))
        _raise_kept(box)
    else:
        text = _text(source)
        st = lib.cusbf_insert_fastx_buffer(filt.handle, text, len(text), b"<stream>",
                                           fill_fraction, stream_handle(stream), ctypes.byref(rep))
    check(lib, st)
    return _report_ins(rep)


def insert_fastx_file(filt: Filter, path, fill_fraction: float = 0.7, stream=None) -> InsertReport:
    """filter.cuh:437-453 (gzip detected by magic bytes)."""
    lib = _lib.load()
    rep = _lib.CInsertReport()
    check(lib, lib.cusbf_insert_fastx_file(filt.handle, os.fsencode(path), fill_fraction,
                                           stream_handle(stream), ctypes.byref(rep)))
    return _report_ins(rep)


def _query(filt: Filter, source, consumer, fill_fraction, stream, path=None) -> QueryReport:
    lib = _lib.load()
    rep = _lib.CQueryReport()
    cfn, cbox = _consumer(consumer)
    rbox = None
    if path is not None:
        st = lib.cusbf_query_fastx_file(filt.handle, os.fsencode(path), fill_fraction,
                                        stream_handle(stream), cfn, None, ctypes.byref(rep))
    elif hasattr(source, "read"):
        rfn, rbox = _reader(source)
        st = lib.cusbf_query_fastx_reader(filt.handle, rfn, None, b"<stream>", fill_fraction,
                                          stream_handle(stream), cfn, None, ctypes.byref(rep))
    else:
        text = _text(source)
        st = lib.cusbf_query_fastx_buffer(filt.handle, text, len(text), b"<stream>", fill_fraction,
                                          stream_handle(stream), cfn, None, ctypes.byref(rep))
    _raise_kept(rbox)
    _raise_kept(cbox)
    check(lib, st)
    return _report_q(rep)


def query_fastx(filt: Filter, source, fill_fraction: float = 0.7, stream=None) -> QueryReport:
    """filter.cuh:632-638: aggregate counts."""
    return _query(filt, source, None, fill_fraction, stream)


def query_fastx_file(filt: Filter, path, fill_fraction: float = 0.7, stream=None) -> QueryReport:
    """filter.cuh:652-668."""
    return _query(filt, None, None, fill_fraction, stream, path=path)


def query_fastx_records(filt: Filter, source, consumer: Callable[[FastxRecordView], None],
                        fill_fraction: float = 0.7, stream=None) -> QueryReport:
    """filter.cuh:683-700: `consumer` gets one FastxRecordView per record, in order."""
    return _query(filt, source, consumer, fill_fraction, stream)


def query_fastx_file_records(filt: Filter, path, consumer: Callable[[FastxRecordView], None],
                             fill_fraction: float = 0.7, stream=None) -> QueryReport:
    """filter.cuh:711-730."""
    return _query(filt, None, consumer, fill_fraction, stream, path=path)


def _detailed(summary_fn) -> FastxDetailedQueryReport:
    records: List[FastxDetailedQueryRecord] = []

    def keep(v: FastxRecordView):
        records.append(FastxDetailedQueryRecord(v.record_index, v.header, v.sequence,
                                                v.queriedBases, v.queriedKmers, v.positive_kmers,
                                                list(v.hits)))

    summary = summary_fn(keep)
    return FastxDetailedQueryReport(summary, records)


def query_fastx_detailed(filt: Filter, source, fill_fraction: float = 0.7,
                         stream=None) -> FastxDetailedQueryReport:
    """filter.cuh:750-760: summary plus one owning record per input record."""
    return _detailed(lambda keep: query_fastx_records(filt, source, keep, fill_fraction, stream))


def query_fastx_file_detailed(filt: Filter, path, fill_fraction: float = 0.7,
                              stream=None) -> FastxDetailedQueryReport:
    """filter.cuh:771-785."""
    return _detailed(lambda keep: query_fastx_file_records(filt, path, keep, fill_fraction,
                                                           stream))


def parse_fastx(source, source_name: str = "<stream>", chunk_bytes: int | None = None,
                max_records: int = 0) -> list:
    """The library's parser alone (no device work): [(header, sequence bytes), ...].
    With chunk_bytes, the records are grouped as the FASTX pipeline groups a text in
    memory (chunks of about chunk_bytes of text, at most max_records records each)."""
    lib = _lib.load()
    out = []

    def keep(v: FastxRecordView):
        out.append((v.header, v.sequence))

    cfn, cbox = _consumer(keep)
    text = _text(source)
    n = ctypes.c_uint64()
    if chunk_bytes is None:
        st = lib.cusbf_parse_fastx_buffer(text, len(text), source_name.encode(), cfn, None,
                                          ctypes.byref(n))
    else:
        st = lib.cusbf_parse_fastx_buffer_chunked(text, len(text), source_name.encode(),
                                                  int(chunk_bytes), int(max_records), cfn, None,
                                                  ctypes.byref(n))
    _raise_kept(cbox)
    check(lib, st)
    return out
