"""Batched record form: many records of uneven length in one persistent launch.

Records are concatenated back to back in one device buffer (no separators)
with a device array of num_records + 1 symbol offsets. Every window inside one
record is inserted / answered exactly as insert_sequence / contains_sequence
would on that record alone (reference: include/cusbf/filter.cuh:179-190,
464-474); a window across a record start is skipped on insert and answered 0,
as the reference answers a window across a separator of a normalised record
batch (normalized_record_batch.hpp:208-232). Query output: one byte per window
start of the concatenation; record r's answers are out[off[r] : off[r+1]-k+1].

`BatchPlan` captures the reset, prepass and persistent kernel of one call into
a hipGraph once and replays it with a single launch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib
from .errors import check
from .filter import Filter, dev_ptr, nbytes, stream_handle

INSERT = 0
CONTAINS = 1
INSERT_SORTED = 2  # table entry of a batched insert on the sorted path (not a plan op)
BUCKETS = 4
BUCKET_BOUNDS = (256, 4096, 65536)  # mean record length (symbols) bounds of buckets 0..2


@dataclass(frozen=True)
class BatchTuning:
    """Launch shape of the batched lane-strip kernels (cusbf_batch_tuning)."""

    waves_per_simd: int = 4
    blocks_per_cu: int = 4
    claim: int = 2  # tiles per work-queue claim; 0 = static striding

    def c(self) -> _lib.CBatchTuning:
        return _lib.CBatchTuning(self.waves_per_simd, self.blocks_per_cu, self.claim, 0)

    @staticmethod
    def from_c(t: _lib.CBatchTuning) -> "BatchTuning":
        return BatchTuning(int(t.waves_per_simd), int(t.blocks_per_cu), int(t.claim))


def batch_variants(filt: Filter) -> tuple:
    """Waves-per-SIMD variants of the filter's batched lane-strip kernels (empty: generic)."""
    v = int(_lib.load().cusbf_batch_variants(filt.handle))
    return tuple(w for bit, w in ((1, 3), (2, 4)) if v & bit)


def batch_bucket(num_symbols: int, num_records: int) -> int:
    return int(_lib.load().cusbf_batch_bucket(int(num_symbols), int(num_records)))


def set_tuning(filt: Filter, bucket: int, op: int, tuning: BatchTuning) -> None:
    lib = _lib.load()
    t = tuning.c()
    check(lib, lib.cusbf_batch_tuning_set(filt.handle, int(bucket), int(op), ctypes.byref(t)))


def get_tuning(filt: Filter, bucket: int, op: int) -> BatchTuning:
    lib = _lib.load()
    t = _lib.CBatchTuning()
    check(lib, lib.cusbf_batch_tuning_get(filt.handle, int(bucket), int(op), ctypes.byref(t)))
    return BatchTuning.from_c(t)


def _symbols(filt: Filter, d_symbols, num_symbols):
    if num_symbols is not None:
        return int(num_symbols)
    return nbytes(d_symbols) // filt.config.symbol_width


def _records(d_offsets, num_records):
    if num_records is not None:
        return int(num_records)
    return d_offsets.numel() - 1


def insert_batched_async(filt: Filter, d_symbols, d_offsets, num_symbols: int | None = None,
                         num_records: int | None = None, stream=None) -> None:
    lib = _lib.load()
    check(lib, lib.cusbf_insert_batched_async(
        filt.handle, dev_ptr(d_symbols), _symbols(filt, d_symbols, num_symbols), dev_ptr(d_offsets),
        _records(d_offsets, num_records), stream_handle(stream)))


def contains_batched_async(filt: Filter, d_symbols, d_offsets, d_output,
                           num_symbols: int | None = None, num_records: int | None = None,
                           output_size: int | None = None, stream=None) -> None:
    lib = _lib.load()
    on = nbytes(d_output) if output_size is None else int(output_size)
    check(lib, lib.cusbf_contains_batched_async(
        filt.handle, dev_ptr(d_symbols), _symbols(filt, d_symbols, num_symbols), dev_ptr(d_offsets),
        _records(d_offsets, num_records), dev_ptr(d_output), on, stream_handle(stream)))


def join_records_async(filt: Filter, d_symbols, d_offsets, d_out=None, num_symbols: int | None = None,
                       num_records: int | None = None, stream=None):
    """The batched records joined by one separator symbol each, on the device (the
    record-batch layout): returns d_out, allocated when not given."""
    import torch

    lib = _lib.load()
    n_sym = _symbols(filt, d_symbols, num_symbols)
    n_rec = _records(d_offsets, num_records)
    out_len = (n_sym + max(n_rec, 1) - 1) * filt.config.symbol_width
    if d_out is None:
        d_out = torch.empty(max(1, out_len), dtype=torch.uint8, device="cuda")[:out_len]
    check(lib, lib.cusbf_join_records_async(filt.handle, dev_ptr(d_symbols), n_sym, dev_ptr(d_offsets), n_rec,
                                            dev_ptr(d_out), nbytes(d_out), stream_handle(stream)))
    return d_out


def batched_status(filt: Filter, stream=None) -> None:
    """Synchronises; raises CusbfError(invalid_argument) for malformed offsets."""
    lib = _lib.load()
    check(lib, lib.cusbf_batched_status(filt.handle, stream_handle(stream)))


def count_batched_windows_runs(filt: Filter, d_symbols, d_offsets, num_symbols: int | None = None,
                               num_records: int | None = None) -> tuple:
    """(valid windows, shard runs) of a batched input."""
    import torch

    lib = _lib.load()
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    check(lib, lib.cusbf_count_batched_windows_runs_async(
        filt.handle, dev_ptr(d_symbols), _symbols(filt, d_symbols, num_symbols), dev_ptr(d_offsets),
        _records(d_offsets, num_records), dev_ptr(counts), 0))
    c = counts.cpu().tolist()
    return int(c[0]), int(c[1])


class BatchPlan:
    """One batched insert or query captured into a hipGraph over fixed buffers."""

    def __init__(self, filt: Filter, op: int, d_symbols, d_offsets, d_output=None,
                 num_symbols: int | None = None, num_records: int | None = None,
                 tuning: BatchTuning | None = None):
        self._lib = _lib.load()
        self.filter = filt
        self.op = op
        # keep the buffers alive as long as the plan
        self._keep = (d_symbols, d_offsets, d_output)
        h = ctypes.c_void_p()
        t = tuning.c() if tuning is not None else None
        on = 0 if d_output is None else nbytes(d_output)
        check(self._lib, self._lib.cusbf_batch_plan_create(
            filt.handle, int(op), dev_ptr(d_symbols), _symbols(filt, d_symbols, num_symbols),
            dev_ptr(d_offsets), _records(d_offsets, num_records), dev_ptr(d_output), on,
            ctypes.byref(t) if t is not None else None, ctypes.byref(h)))
        self._h = h

    def launch(self, stream=None) -> None:
        check(self._lib, self._lib.cusbf_batch_plan_launch(self._h, stream_handle(stream)))

    def status(self, stream=None) -> None:
        check(self._lib, self._lib.cusbf_batch_plan_status(self._h, stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cusbf_batch_plan_destroy(self._h)
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
