"""Filter: Python mirror of cusbf::filter<Config> (reference: include/cusbf/filter.cuh:68-1848).

Every method calls the HIP library through the C ABI; device memory is any
object exposing a device address (torch tensors on the GPU, or plain ints),
streams are torch.cuda.Stream objects, raw hipStream_t ints, or None for the
default stream. Host variants synchronise before returning, *_async variants
do not, exactly as in the reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Iterable, Sequence, Tuple

import numpy as np

from . import _lib
from .config import Config
from .errors import CusbfError, ErrorCategory, check


@dataclass
class RecordRange:
    """Fastx.hpp:24-29"""

    sequence_offset: int
    sequence_bytes: int


@dataclass
class InsertReport:
    """FastxInsertReport (Fastx.hpp:150-158)."""

    records_indexed: int = 0
    indexed_bases: int = 0
    inserted_kmers: int = 0


@dataclass
class QueryReport:
    """FastxQueryReport (Fastx.hpp:161-170)."""

    records_queried: int = 0
    queried_bases: int = 0
    queried_kmers: int = 0
    positive_kmers: int = 0


def dev_ptr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if hasattr(x, "is_cuda") and not x.is_cuda:
            raise CusbfError(ErrorCategory.invalid_argument, "expected a device tensor")
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device address of {type(x)!r}")


def nbytes(x) -> int:
    if hasattr(x, "numel") and hasattr(x, "element_size"):
        return int(x.numel() * x.element_size())
    raise TypeError("pass the size explicitly for raw device pointers")


def stream_handle(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    if hasattr(stream, "cuda_stream"):
        return int(stream.cuda_stream)
    raise TypeError(f"unsupported stream object {type(stream)!r}")


def _host_bytes(seq) -> np.ndarray:
    if isinstance(seq, str):
        seq = seq.encode("latin-1")
    if isinstance(seq, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(seq), dtype=np.uint8)
    return np.ascontiguousarray(seq, dtype=np.uint8)


def _ranges(ranges: Iterable) -> Tuple[ctypes.Array, int]:
    items = [r if isinstance(r, RecordRange) else RecordRange(int(r[0]), int(r[1])) for r in ranges]
    arr = (_lib.CRecordRange * max(1, len(items)))()
    for i, r in enumerate(items):
        arr[i].sequence_offset = r.sequence_offset
        arr[i].sequence_bytes = r.sequence_bytes
    return arr, len(items)


class _DeviceArray:
    """__cuda_array_interface__ description of library-owned device memory."""

    def __init__(self, ptr: int, n: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "version": 3, "strides": None}


# Insert paths of the lane-strip kernels (include/cusbf_hip.h, CUSBF_INSERT_*).
INSERT_AUTO = 0
INSERT_ATOMIC = 1
INSERT_SORTED = 2


@dataclass
class InsertPathInfo:
    path: int
    slice_bits: int
    reserve_per_tile: int
    last_path: int  # INSERT_ATOMIC or INSERT_SORTED: the path the last insert call took


@dataclass
class SortedWorkspaceInfo:
    limit: int  # bytes (0: half of the free device memory)
    allocated: int  # bytes the workspace holds now
    last_chunks: int  # sorted inserts the last plain insert ran as (0: atomic path)


class Filter:
    """GPU-resident Super Bloom filter (filter.cuh:68-1848)."""

    def __init__(self, config: Config, requested_filter_bits: int):
        self.config = config
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib, self._lib.cusbf_create(ctypes.byref(config.c_config),
                                                int(requested_filter_bits), ctypes.byref(h)))
        self._h = h

    # lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cusbf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    # filter images (no reference counterpart) -------------------------------------
    def save(self, path) -> None:
        """Writes the filter image (header + words) to `path`."""
        import os

        check(self._lib, self._lib.cusbf_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path) -> "Filter":
        """A filter recreated from an image on the current device (checksum verified)."""
        import os

        lib = _lib.load()
        c = _lib.CConfig()
        bits = ctypes.c_uint64()
        check(lib, lib.cusbf_image_config(os.fsencode(path), ctypes.byref(c), ctypes.byref(bits)))
        config = Config.from_c(c)
        h = ctypes.c_void_p()
        check(lib, lib.cusbf_load(os.fsencode(path), ctypes.byref(h)))
        f = cls.__new__(cls)
        f.config, f._lib, f._h = config, lib, h
        return f

    # properties (filter.cuh:391-414) ---------------------------------------
    def filter_bits(self) -> int:
        return int(self._lib.cusbf_filter_bits(self._h))

    def num_shards(self) -> int:
        return int(self._lib.cusbf_num_shards(self._h))

    def load_factor(self) -> float:
        out = ctypes.c_float()
        check(self._lib, self._lib.cusbf_load_factor(self._h, ctypes.byref(out)))
        return float(out.value)

    def clear(self, stream=None) -> None:
        check(self._lib, self._lib.cusbf_clear(self._h, stream_handle(stream)))

    def kernel_family(self) -> str:
        return self._lib.cusbf_kernel_family(self._h).decode()

    # insert path (no reference counterpart) ------------------------------------
    def set_insert_path(self, path: int, slice_bits: int = 0, reserve_per_tile: int = 0) -> None:
        """INSERT_AUTO / INSERT_ATOMIC / INSERT_SORTED for the lane-strip inserts;
        slice_bits (1..12, 0 = 12) and records reserved per wave tile (1..1024, 0 =
        from the minimizer span) shape the sorted path."""
        check(self._lib, self._lib.cusbf_set_insert_path(self._h, int(path), int(slice_bits),
                                                         int(reserve_per_tile)))

    def insert_path(self) -> InsertPathInfo:
        v = [ctypes.c_uint32() for _ in range(4)]
        check(self._lib, self._lib.cusbf_get_insert_path(self._h, *[ctypes.byref(x) for x in v]))
        return InsertPathInfo(*(int(x.value) for x in v))

    def set_sorted_workspace_limit(self, nbytes: int) -> None:
        """Bound the sorted insert's workspace (0: half of the free device memory);
        a plain insert whose records do not fit runs as several sorted inserts over
        consecutive window ranges."""
        check(self._lib, self._lib.cusbf_set_sorted_workspace_limit(self._h, int(nbytes)))

    def sorted_workspace(self) -> SortedWorkspaceInfo:
        lim, alloc, chunks = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        check(self._lib, self._lib.cusbf_get_sorted_workspace(self._h, ctypes.byref(lim), ctypes.byref(alloc),
                                                              ctypes.byref(chunks)))
        return SortedWorkspaceInfo(int(lim.value), int(alloc.value), int(chunks.value))

    def words_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(self._lib, self._lib.cusbf_words_device_ptr(self._h, ctypes.byref(p)))
        return int(p.value or 0)

    def words_tensor(self):
        """The filter words as an int64 torch tensor viewing the library's device
        memory (no copy; valid while the filter lives)."""
        import torch

        view = _DeviceArray(self.words_device_ptr(), self.num_shards() * 4, "<i8")
        return torch.as_tensor(view, device="cuda")

    def words(self) -> np.ndarray:
        """All num_shards*4 filter words (synchronous copy to the host)."""
        out = np.zeros(self.num_shards() * 4, dtype=np.uint64)
        check(self._lib, self._lib.cusbf_copy_words_to_host(self._h, out.ctypes.data, out.size))
        return out

    # counts ----------------------------------------------------------------
    def record_symbol_count(self, nbytes_: int) -> int:
        return nbytes_ // self.config.symbol_width

    def record_kmer_count(self, nbytes_: int) -> int:
        s = self.record_symbol_count(nbytes_)
        return 0 if s < self.config.k else s - self.config.k + 1

    def dense_packed_word_count(self, num_symbols: int) -> int:
        return int(self._lib.cusbf_dense_packed_word_count(ctypes.byref(self.config.c_config),
                                                           num_symbols))

    def dense_packed_kmer_count(self, num_symbols: int) -> int:
        return int(self._lib.cusbf_dense_packed_kmer_count(ctypes.byref(self.config.c_config),
                                                           num_symbols))

    # byte sequences ----------------------------------------------------------
    def insert_sequence(self, sequence, stream=None) -> int:
        """filter.cuh:154-167: host bytes; returns all k-mer windows."""
        b = _host_bytes(sequence)
        out = ctypes.c_uint64()
        check(self._lib, self._lib.cusbf_insert_sequence(self._h, b.ctypes.data, b.size,
                                                         stream_handle(stream), ctypes.byref(out)))
        return int(out.value)

    def insert_sequence_async(self, d_sequence, size: int | None = None, stream=None) -> int:
        """filter.cuh:179-190"""
        n = nbytes(d_sequence) if size is None else int(size)
        out = ctypes.c_uint64()
        check(self._lib, self._lib.cusbf_insert_sequence_async(
            self._h, dev_ptr(d_sequence), n, stream_handle(stream), ctypes.byref(out)))
        return int(out.value)

    def contains_sequence(self, sequence, d_output=None, stream=None) -> np.ndarray | None:
        """filter.cuh:514-533: host bytes in, one byte per window out. With a device
        d_output (filter.cuh:486-501) the answers are written there instead and
        nothing is returned."""
        b = _host_bytes(sequence)
        if d_output is not None:
            check(self._lib, self._lib.cusbf_contains_sequence_host_to_device(
                self._h, b.ctypes.data, b.size, dev_ptr(d_output), nbytes(d_output),
                stream_handle(stream)))
            return None
        out = np.zeros(self.record_kmer_count(b.size), dtype=np.uint8)
        if out.size == 0:
            return out
        check(self._lib, self._lib.cusbf_contains_sequence(self._h, b.ctypes.data, b.size,
                                                           out.ctypes.data, out.size,
                                                           stream_handle(stream)))
        return out

    def contains_sequence_async(self, d_sequence, d_output, size: int | None = None,
                                output_size: int | None = None, stream=None) -> None:
        """filter.cuh:464-474"""
        n = nbytes(d_sequence) if size is None else int(size)
        on = (nbytes(d_output) if output_size is None else int(output_size))
        check(self._lib, self._lib.cusbf_contains_sequence_async(
            self._h, dev_ptr(d_sequence), n, dev_ptr(d_output), on, stream_handle(stream)))

    # dense packed ---------------------------------------------------------------
    def insert_dense_packed_async(self, d_words, num_symbols: int, num_words: int | None = None,
                                  stream=None) -> int:
        """filter.cuh:207-219"""
        nw = nbytes(d_words) // 8 if num_words is None else int(num_words)
        out = ctypes.c_uint64()
        check(self._lib, self._lib.cusbf_insert_dense_packed_async(
            self._h, dev_ptr(d_words), nw, int(num_symbols), stream_handle(stream),
            ctypes.byref(out)))
        return int(out.value)

    def contains_dense_packed_async(self, d_words, num_symbols: int, d_output,
                                    num_words: int | None = None,
                                    output_size: int | None = None, stream=None) -> None:
        """filter.cuh:248-259"""
        nw = nbytes(d_words) // 8 if num_words is None else int(num_words)
        on = nbytes(d_output) if output_size is None else int(output_size)
        check(self._lib, self._lib.cusbf_contains_dense_packed_async(
            self._h, dev_ptr(d_words), nw, int(num_symbols), dev_ptr(d_output), on,
            stream_handle(stream)))

    def insert_dense_packed(self, words: Sequence[int] | np.ndarray, num_symbols: int,
                            stream=None) -> int:
        """filter.cuh:226-240 (host words, synchronises)."""
        w = np.ascontiguousarray(words, dtype=np.uint64)
        inserted = ctypes.c_uint64(0)
        check(self._lib, self._lib.cusbf_insert_dense_packed(
            self._h, w.ctypes.data, w.size, num_symbols, stream_handle(stream), ctypes.byref(inserted)))
        return int(inserted.value)

    def contains_dense_packed(self, words: Sequence[int] | np.ndarray, num_symbols: int,
                              d_output=None, stream=None) -> np.ndarray | None:
        """filter.cuh:289-312 (host words in, host hits out); with a device d_output,
        filter.cuh:267-282 (answers written there, nothing returned)."""
        if d_output is not None:
            w = np.ascontiguousarray(words, dtype=np.uint64)
            check(self._lib, self._lib.cusbf_contains_dense_packed_host_to_device(
                self._h, w.ctypes.data, w.size, int(num_symbols), dev_ptr(d_output),
                nbytes(d_output), stream_handle(stream)))
            return None
        n = self.dense_packed_kmer_count(num_symbols)
        out = np.zeros(n, dtype=np.uint8)
        if n == 0:
            return out
        w = np.ascontiguousarray(words, dtype=np.uint64)
        check(self._lib, self._lib.cusbf_contains_dense_packed(
            self._h, w.ctypes.data, w.size, num_symbols, out.ctypes.data, out.size,
            stream_handle(stream)))
        return out

    # record batches -------------------------------------------------------------
    def insert_record_batch(self, sequence, ranges: Iterable, stream=None) -> InsertReport:
        """filter.cuh:338-364 (host batch, separators injected)."""
        b = _host_bytes(sequence)
        arr, n = _ranges(ranges)
        rep = _lib.CInsertReport()
        check(self._lib, self._lib.cusbf_insert_record_batch(
            self._h, b.ctypes.data, b.size, arr, n, stream_handle(stream), ctypes.byref(rep)))
        return InsertReport(rep.records_indexed, rep.indexed_bases, rep.inserted_kmers)

    def query_record_batch(self, sequence, ranges: Iterable, stream=None) -> QueryReport:
        """filter.cuh:549-552, 930-990 (aggregate form)."""
        b = _host_bytes(sequence)
        arr, n = _ranges(ranges)
        rep = _lib.CQueryReport()
        check(self._lib, self._lib.cusbf_query_record_batch(
            self._h, b.ctypes.data, b.size, arr, n, stream_handle(stream), ctypes.byref(rep)))
        return QueryReport(rep.records_queried, rep.queried_bases, rep.queried_kmers,
                           rep.positive_kmers)

    def insert_record_batch_async(self, d_sequence, ranges: Iterable, size: int | None = None,
                                  stream=None) -> int:
        """filter.cuh:385-400 (device batch with separators embedded)."""
        n = nbytes(d_sequence) if size is None else int(size)
        arr, nr = _ranges(ranges)
        out = ctypes.c_uint64()
        check(self._lib, self._lib.cusbf_insert_record_batch_async(
            self._h, dev_ptr(d_sequence), n, arr, nr, stream_handle(stream), ctypes.byref(out)))
        return int(out.value)

    def contains_record_batch_async(self, d_sequence, ranges: Iterable, d_output,
                                    size: int | None = None, output_size: int | None = None,
                                    stream=None) -> None:
        """filter.cuh:603-618"""
        n = nbytes(d_sequence) if size is None else int(size)
        on = nbytes(d_output) if output_size is None else int(output_size)
        arr, nr = _ranges(ranges)
        check(self._lib, self._lib.cusbf_contains_record_batch_async(
            self._h, dev_ptr(d_sequence), n, arr, nr, dev_ptr(d_output), on,
            stream_handle(stream)))

    # FASTA / FASTQ (filter.cuh:402-453, 620-800); see cusbf_amd.fastx ------------
    def insert_fastx(self, source, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.insert_fastx(self, source, fill_fraction, stream)

    def insert_fastx_file(self, path, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.insert_fastx_file(self, path, fill_fraction, stream)

    def query_fastx(self, source, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx(self, source, fill_fraction, stream)

    def query_fastx_file(self, path, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx_file(self, path, fill_fraction, stream)

    def query_fastx_records(self, source, consumer, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx_records(self, source, consumer, fill_fraction, stream)

    def query_fastx_file_records(self, path, consumer, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx_file_records(self, path, consumer, fill_fraction, stream)

    def query_fastx_detailed(self, source, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx_detailed(self, source, fill_fraction, stream)

    def query_fastx_file_detailed(self, path, fill_fraction: float = 0.7, stream=None):
        from . import fastx
        return fastx.query_fastx_file_detailed(self, path, fill_fraction, stream)

    # diagnostics -----------------------------------------------------------------
    def count_windows_runs(self, d_sequence, size: int | None = None) -> Tuple[int, int]:
        """(valid windows, shard runs) of a device sequence (byte model of the roofline)."""
        import torch

        n = nbytes(d_sequence) if size is None else int(size)
        counts = torch.zeros(2, dtype=torch.int64, device="cuda")
        check(self._lib, self._lib.cusbf_count_windows_runs_async(
            self._h, dev_ptr(d_sequence), n, dev_ptr(counts), 0))
        c = counts.cpu().tolist()
        return int(c[0]), int(c[1])


def pack_dense_sequence(config: Config, sequence) -> np.ndarray:
    """Host packer (dense_packed.hpp:35-49): LSB-first, 64/symbol_bits symbols per word."""
    b = _host_bytes(sequence)
    w = config.symbol_width
    nsym = b.size // w
    per = 64 // config.symbol_bits
    out = np.zeros((nsym + per - 1) // per, dtype=np.uint64)
    if nsym == 0:
        return out
    codes = encode_symbols(config, b[: nsym * w])
    codes = codes.astype(np.uint64) & np.uint64(config.symbol_mask)
    idx = np.arange(nsym, dtype=np.uint64)
    shifts = (idx % np.uint64(per)) * np.uint64(config.symbol_bits)
    np.bitwise_or.at(out, (idx // np.uint64(per)).astype(np.int64), codes << shifts)
    return out


def encode_symbols(config: Config, b: np.ndarray) -> np.ndarray:
    """Vectorised Alphabet::encode over a byte array (Alphabet.cuh:132-217)."""
    a = config.alphabet
    if a.code == 0:
        up = b & 0xDF
        valid = (up == ord("A")) | (up == ord("C")) | (up == ord("G")) | (up == ord("T"))
        return np.where(valid, (b >> 1) & 3, 0xFF).astype(np.uint8)
    if a.code == 1:
        t = b.reshape(-1, 3)
        e = [encode_symbols(Config(3, 2, 2), t[:, i].copy()) for i in range(3)]
        valid = (e[0] != 0xFF) & (e[1] != 0xFF) & (e[2] != 0xFF)
        return np.where(valid, (e[0] << 4) | (e[1] << 2) | e[2], 0xFF).astype(np.uint8)
    if a.code == 2:
        idx = ((b & 0xDF).astype(np.int16) - ord("A")).astype(np.uint8)
        return np.where(idx < 26, idx, 0xFF).astype(np.uint8)
    lut = np.frombuffer(a.lut(), dtype=np.uint8)
    return lut[b]


def pack_dense_sequence_device(config: Config, d_sequence, num_symbols: int, d_words,
                               stream=None, num_words: int | None = None) -> None:
    """dense_packed.hpp:57-73"""
    lib = _lib.load()
    nw = nbytes(d_words) // 8 if num_words is None else int(num_words)
    check(lib, lib.cusbf_pack_dense_sequence_async(ctypes.byref(config.c_config),
                                                   dev_ptr(d_sequence), int(num_symbols),
                                                   dev_ptr(d_words), nw, stream_handle(stream)))


def count_positive(d_hits, num_hits: int | None = None, stream=None) -> int:
    """count_positive_kmers_total (count_positive_kmers.cuh:47-62)."""
    import torch

    lib = _lib.load()
    n = nbytes(d_hits) if num_hits is None else int(num_hits)
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    check(lib, lib.cusbf_count_positive_async(dev_ptr(d_hits), n, dev_ptr(out),
                                              stream_handle(stream)))
    return int(out.item())
