"""ctypes binding of the CPU oracle (oracle/sbf_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against. Only
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import
this module; the product package never does.

The naive functions restate the reference's per-window arithmetic (see the
file:line citations in sbf_oracle.c); the `mt_*` functions are the fast
multithreaded form used as the CPU Super Bloom baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsbf_oracle.so")

ALPHA_DNA = 0
ALPHA_DNA_TRIPLET = 1
ALPHA_PROTEIN = 2
ALPHA_CUSTOM = 3


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_uint32),
        ("s", ctypes.c_uint32),
        ("m", ctypes.c_uint32),
        ("hash_count", ctypes.c_uint32),
        ("alphabet", ctypes.c_uint32),
        ("symbol_width", ctypes.c_uint32),
        ("symbol_count", ctypes.c_uint32),
        ("symbol_bits", ctypes.c_uint32),
        ("invalid_symbol", ctypes.c_uint8),
        ("separator", ctypes.c_uint8),
        ("custom_lut", ctypes.c_uint8 * 256),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle shared library with its Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        args = ["make", "-C", _HERE]
        if force:
            args.append("-B")
        subprocess.run(args, check=True, capture_output=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        u64 = ctypes.c_uint64
        u32 = ctypes.c_uint32
        cfgp = ctypes.POINTER(OrcConfig)
        L.orc_config_init.argtypes = [cfgp, u32, u32, u32, u32, u32, P, u32, ctypes.c_uint8,
                                      ctypes.c_uint8]
        L.orc_config_init.restype = ctypes.c_int
        L.orc_encode.argtypes = [cfgp, ctypes.c_char_p]
        L.orc_encode.restype = ctypes.c_uint8
        L.orc_hash64.argtypes = [u64]
        L.orc_hash64.restype = u64
        L.orc_minimizer_hash64.argtypes = [u64]
        L.orc_minimizer_hash64.restype = u64
        L.orc_salt.argtypes = [u32]
        L.orc_salt.restype = u64
        L.orc_bit_address.argtypes = [cfgp, u32, u64]
        L.orc_bit_address.restype = u32
        L.orc_num_shards.argtypes = [u64]
        L.orc_num_shards.restype = u64
        L.orc_kmer_count.argtypes = [cfgp, u64]
        L.orc_kmer_count.restype = u64
        L.orc_window_valid.argtypes = [cfgp, P, u64]
        L.orc_window_valid.restype = ctypes.c_int
        L.orc_pack_window.argtypes = [cfgp, P, u64]
        L.orc_pack_window.restype = u64
        L.orc_window_minimizer_hash.argtypes = [cfgp, u64]
        L.orc_window_minimizer_hash.restype = u64
        L.orc_shard_index.argtypes = [u64, u64]
        L.orc_shard_index.restype = u32
        L.orc_window_masks.argtypes = [cfgp, u64, P]
        L.orc_window_masks.restype = None
        for name in ("orc_insert", "orc_mt_insert"):
            fn = getattr(L, name)
            fn.argtypes = [cfgp, P, u64, P, u64] + ([ctypes.c_int] if "mt" in name else [])
            fn.restype = None
        for name in ("orc_contains", "orc_mt_contains"):
            fn = getattr(L, name)
            fn.argtypes = [cfgp, P, u64, P, u64, P] + ([ctypes.c_int] if "mt" in name else [])
            fn.restype = None
        L.orc_dense_word_count.argtypes = [cfgp, u64]
        L.orc_dense_word_count.restype = u64
        L.orc_pack_dense.argtypes = [cfgp, P, u64, P]
        L.orc_pack_dense.restype = None
        L.orc_insert_dense.argtypes = [cfgp, P, u64, P, u64]
        L.orc_insert_dense.restype = None
        L.orc_contains_dense.argtypes = [cfgp, P, u64, P, u64, P]
        L.orc_contains_dense.restype = None
        L.orc_count.argtypes = [cfgp, P, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.orc_count.restype = None
        L.orc_hardware_threads.argtypes = []
        L.orc_hardware_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _buf(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray)):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    if isinstance(data, str):
        return np.frombuffer(data.encode("latin-1"), dtype=np.uint8)
    return np.ascontiguousarray(data, dtype=np.uint8)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


class Oracle:
    """One oracle configuration (k, s, m, hash count, alphabet)."""

    def __init__(self, k, s, m, hash_count=4, alphabet=ALPHA_DNA, custom_lut=None,
                 custom_symbol_count=0, custom_invalid=0xFF, custom_separator=0):
        self.cfg = OrcConfig()
        lut = None
        if custom_lut is not None:
            lut_arr = (ctypes.c_uint8 * 256)(*custom_lut)
            lut = ctypes.cast(lut_arr, ctypes.c_void_p)
        rc = lib().orc_config_init(ctypes.byref(self.cfg), k, s, m, hash_count, alphabet, lut,
                                   custom_symbol_count, custom_invalid, custom_separator)
        if rc != 0:
            raise ValueError("invalid oracle configuration")

    # scalar helpers -------------------------------------------------------
    @property
    def symbol_width(self):
        return self.cfg.symbol_width

    @property
    def symbol_bits(self):
        return self.cfg.symbol_bits

    @property
    def separator(self):
        return self.cfg.separator

    def encode(self, data) -> int:
        return lib().orc_encode(ctypes.byref(self.cfg), bytes(data))

    def kmer_count(self, nbytes: int) -> int:
        return lib().orc_kmer_count(ctypes.byref(self.cfg), nbytes)

    @staticmethod
    def num_shards(requested_bits: int) -> int:
        return lib().orc_num_shards(requested_bits)

    def new_filter(self, num_shards: int) -> np.ndarray:
        return np.zeros(num_shards * 4, dtype=np.uint64)

    # naive form ----------------------------------------------------------
    def insert(self, seq, words: np.ndarray) -> None:
        b = _buf(seq)
        lib().orc_insert(ctypes.byref(self.cfg), _ptr(b), b.size, _ptr(words), words.size // 4)

    def contains(self, seq, words: np.ndarray) -> np.ndarray:
        b = _buf(seq)
        out = np.zeros(self.kmer_count(b.size), dtype=np.uint8)
        lib().orc_contains(ctypes.byref(self.cfg), _ptr(b), b.size, _ptr(words), words.size // 4,
                           _ptr(out))
        return out

    def pack_dense(self, seq) -> np.ndarray:
        b = _buf(seq)
        nsym = b.size // self.symbol_width
        out = np.zeros(lib().orc_dense_word_count(ctypes.byref(self.cfg), nsym), dtype=np.uint64)
        lib().orc_pack_dense(ctypes.byref(self.cfg), _ptr(b), nsym, _ptr(out))
        return out

    def insert_dense(self, packed: np.ndarray, num_symbols: int, words: np.ndarray) -> None:
        lib().orc_insert_dense(ctypes.byref(self.cfg), _ptr(packed), num_symbols, _ptr(words),
                               words.size // 4)

    def contains_dense(self, packed: np.ndarray, num_symbols: int, words: np.ndarray):
        n = 0 if num_symbols < self.cfg.k else num_symbols - self.cfg.k + 1
        out = np.zeros(n, dtype=np.uint8)
        lib().orc_contains_dense(ctypes.byref(self.cfg), _ptr(packed), num_symbols, _ptr(words),
                                 words.size // 4, _ptr(out))
        return out

    def count(self, seq, num_shards: int):
        """(valid windows, shard runs) over the byte path."""
        b = _buf(seq)
        v = ctypes.c_uint64()
        r = ctypes.c_uint64()
        lib().orc_count(ctypes.byref(self.cfg), _ptr(b), b.size, num_shards, ctypes.byref(v),
                        ctypes.byref(r))
        return v.value, r.value

    def window(self, seq, j: int):
        """(valid, packed k-mer, minimizer hash, 4 masks) of window j."""
        b = _buf(seq)
        L = lib()
        valid = L.orc_window_valid(ctypes.byref(self.cfg), _ptr(b), j)
        packed = L.orc_pack_window(ctypes.byref(self.cfg), _ptr(b), j)
        mh = L.orc_window_minimizer_hash(ctypes.byref(self.cfg), packed)
        masks = np.zeros(4, dtype=np.uint64)
        L.orc_window_masks(ctypes.byref(self.cfg), packed, _ptr(masks))
        return bool(valid), packed, mh, masks

    def bit_address(self, hash_index: int, base_hash: int) -> int:
        return lib().orc_bit_address(ctypes.byref(self.cfg), hash_index, base_hash)

    # fast multithreaded form (CPU baseline) -------------------------------
    def mt_insert(self, seq, words: np.ndarray, threads: int = 0) -> None:
        b = _buf(seq)
        lib().orc_mt_insert(ctypes.byref(self.cfg), _ptr(b), b.size, _ptr(words),
                            words.size // 4, threads or default_threads())

    def mt_contains(self, seq, words: np.ndarray, threads: int = 0) -> np.ndarray:
        b = _buf(seq)
        out = np.zeros(self.kmer_count(b.size), dtype=np.uint8)
        lib().orc_mt_contains(ctypes.byref(self.cfg), _ptr(b), b.size, _ptr(words),
                              words.size // 4, _ptr(out), threads or default_threads())
        return out


def hash64(x: int) -> int:
    return lib().orc_hash64(x)


def minimizer_hash64(x: int) -> int:
    return lib().orc_minimizer_hash64(x)


def salt(i: int) -> int:
    return lib().orc_salt(i)


def hardware_threads() -> int:
    return lib().orc_hardware_threads()


def normalize_records(seq, ranges, separator: int, width: int = 1) -> bytes:
    """Records of `seq` joined as a normalised record batch
    (normalized_record_batch.hpp:208-276): before every record after the first
    non-empty output, the running size is padded with separator bytes to a
    multiple of the symbol width and one separator symbol (`width` bytes) is
    written; then the record's bytes."""
    b = _buf(seq)
    parts = []
    size = 0
    sep = bytes([separator])
    for off, n in ranges:
        off, n = int(off), int(n)
        if size:
            pad = (width - size % width) % width + width
            parts.append(sep * pad)
            size += pad
        parts.append(b[off:off + n].tobytes())
        size += n
    return b"".join(parts)


def default_threads() -> int:
    """Worker threads for the multithreaded form: the CPUs this process may run on
    (affinity mask and cgroup CPU quota), at most 16 (one GPU's share of a host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, -(-int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, min(n, 16))
