"""Server: the query path as a continuous-batching request server (C ABI cusbf_server_*).

The filter stays resident on the GPU. `submit` queues a host record batch and
returns a ticket; the library's scheduler admits queued requests into batch
slots as slots free up, and `wait` returns the request's answers (one byte per
window of its normalised record batch, identical to
Filter.query_record_batch_hits on the request alone), its query report and
its latency (submit to answers ready).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Iterable

import numpy as np

from . import _lib
from .errors import check
from .filter import Filter, QueryReport, RecordRange, _host_bytes


def ranges_array(ranges) -> tuple:
    """(ctypes array of cusbf_record_range, count) from (offset, bytes) pairs or an (n, 2) array."""
    if isinstance(ranges, np.ndarray):
        a = np.ascontiguousarray(ranges, dtype=np.uint64).reshape(-1, 2)
        n = a.shape[0]
        arr = (_lib.CRecordRange * max(1, n)).from_buffer_copy(
            a.tobytes() if n else bytes(16))
        return arr, n
    items = [r if isinstance(r, RecordRange) else RecordRange(int(r[0]), int(r[1]))
             for r in ranges]
    arr = (_lib.CRecordRange * max(1, len(items)))()
    for i, r in enumerate(items):
        arr[i].sequence_offset = r.sequence_offset
        arr[i].sequence_bytes = r.sequence_bytes
    return arr, len(items)


def normalized_size(arr, n: int, width: int) -> int:
    """Bytes of a normalised record batch (normalized_record_batch.hpp:208-232)."""
    sizes = np.frombuffer(arr, dtype=np.uint64, count=2 * n)[1::2] if n else np.zeros(0, np.uint64)
    if width == 1:
        nz = np.flatnonzero(sizes)
        seps = 0 if nz.size == 0 else n - int(nz[0]) - 1
        return int(sizes.sum()) + seps
    size = 0
    for b in sizes.tolist():
        if size:
            rem = size % width
            size += (width - rem if rem else 0) + width
        size += int(b)
    return size


@dataclass
class Pending:
    ticket: int
    hits: np.ndarray
    keep: tuple  # host buffers that must outlive the request


class Server:
    def __init__(self, filt: Filter, staging_bytes: int = 1 << 30, slots: int = 4,
                 max_wait_us: int = 0):
        self._lib = _lib.load()
        self.filter = filt
        opts = _lib.CServerOptions(int(staging_bytes), int(slots), int(max_wait_us))
        h = ctypes.c_void_p()
        check(self._lib, self._lib.cusbf_server_create(filt.handle, ctypes.byref(opts),
                                                       ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cusbf_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def request_windows(self, ranges) -> int:
        arr, n = ranges_array(ranges)
        return int(self._lib.cusbf_server_request_windows(self._h, arr, n))

    def submit(self, sequence, ranges, out: np.ndarray | None = None) -> Pending:
        """Queue one request; `out` (uint8, at least the request's window count)
        receives the answers, else a new array is allocated."""
        seq = _host_bytes(sequence)
        arr, n = ranges_array(ranges)
        windows = int(self._lib.cusbf_server_request_windows(self._h, arr, n))
        hits = out if out is not None else np.zeros(max(1, windows), dtype=np.uint8)
        if hits.dtype != np.uint8 or not hits.flags.c_contiguous or hits.size < windows:
            raise ValueError("out must be a contiguous uint8 array holding every window")
        t = ctypes.c_uint64()
        check(self._lib, self._lib.cusbf_server_submit(self._h, seq.ctypes.data, seq.size, arr, n,
                                                       hits.ctypes.data, hits.size,
                                                       ctypes.byref(t)))
        return Pending(int(t.value), hits[:windows], (seq, arr))

    def wait(self, pending: Pending, report: bool = True):
        """(hits, QueryReport or None, latency seconds)."""
        rep = _lib.CQueryReport()
        lat = ctypes.c_double()
        check(self._lib, self._lib.cusbf_server_wait(
            self._h, pending.ticket, ctypes.byref(rep) if report else None, ctypes.byref(lat)))
        q = QueryReport(rep.records_queried, rep.queried_bases, rep.queried_kmers,
                        rep.positive_kmers) if report else None
        return pending.hits, q, float(lat.value)

    def stats(self) -> dict:
        st = _lib.CServerStats()
        check(self._lib, self._lib.cusbf_server_stats_get(self._h, ctypes.byref(st)))
        return {"requests": st.requests, "batches": st.batches,
                "max_requests_per_batch": st.max_requests_per_batch, "bytes": st.bytes,
                "windows": st.windows, "fill_seconds": st.fill_seconds,
                "launch_seconds": st.launch_seconds, "device_seconds": st.device_seconds,
                "copy_seconds": st.copy_seconds}


def query_record_batch_hits(filt: Filter, sequence, ranges, stream=None,
                            out: np.ndarray | None = None):
    """filter.cuh:566-582, 1007-1131: one request alone -> (hits, QueryReport)."""
    from .filter import stream_handle

    lib = _lib.load()
    seq = _host_bytes(sequence)
    arr, n = ranges_array(ranges)
    windows = filt.record_kmer_count(normalized_size(arr, n, filt.config.symbol_width))
    hits = out if out is not None else np.zeros(max(1, windows), dtype=np.uint8)
    if hits.dtype != np.uint8 or not hits.flags.c_contiguous or hits.size < windows:
        raise ValueError("out must be a contiguous uint8 array holding every window")
    rep = _lib.CQueryReport()
    check(lib, lib.cusbf_query_record_batch_hits(filt.handle, seq.ctypes.data, seq.size, arr, n,
                                                  stream_handle(stream), hits.ctypes.data,
                                                  hits.size, ctypes.byref(rep)))
    return hits[:windows], QueryReport(rep.records_queried, rep.queried_bases, rep.queried_kmers,
                                       rep.positive_kmers)
