"""Restatement of the filter image format and its checksum, and of load_factor.

TEST INFRASTRUCTURE ONLY: the checker that the library's filter images and
load factor are compared against. Only tests/ import this module.

Filter images have no counterpart in the reference (it keeps filters only in
device memory); the format is this library's own: a 336-byte little-endian
header (magic "CUSBFIMG", version, header size, the 284-byte cusbf_config,
4 reserved bytes, shard count, filter bits, word count, checksum) followed by
the filter words in device order. The checksum is

    sum over i of splitmix64_mix(words[i] XOR (i * 0x9E3779B97F4A7C15))  mod 2^64

where splitmix64_mix is the finaliser of Vigna's SplitMix64 generator; that
generator's published outputs from seed 0 pin the finaliser here
(tests/test_oracle_pins.py).

load_factor restates filter.cuh:810-823 of the reference: popcount over all
words as a uint64, then float(set bits) / float(filter bits) in IEEE single
precision.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
HEADER_DTYPE = np.dtype([
    ("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"),
    ("config", "V284"), ("reserved", "<u4"),
    ("num_shards", "<u8"), ("filter_bits", "<u8"), ("word_count", "<u8"), ("checksum", "<u8"),
])
assert HEADER_DTYPE.itemsize == 336


def splitmix64_mix(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser, element-wise over uint64 (arithmetic mod 2^64)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix64_sequence(seed: int, n: int) -> list:
    """The first n outputs of SplitMix64 from `seed` (state += golden, output mix(state))."""
    state = np.uint64(seed)
    out = []
    with np.errstate(over="ignore"):
        for _ in range(n):
            state = state + GOLDEN
            out.append(int(splitmix64_mix(np.array([state]))[0]))
    return out


def words_checksum(words: np.ndarray, chunk: int = 1 << 26) -> int:
    """The image checksum of `words` (uint64), in chunks to bound host memory."""
    w = np.asarray(words, dtype=np.uint64).reshape(-1)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for a in range(0, w.size, chunk):
            idx = np.arange(a, min(w.size, a + chunk), dtype=np.uint64)
            total = total + splitmix64_mix(w[a:a + idx.size] ^ (idx * GOLDEN)).sum(dtype=np.uint64)
    return int(total)


def read_image(path: str):
    """(header record, words) of a filter image file."""
    raw = np.fromfile(path, dtype=np.uint8)
    header = raw[:HEADER_DTYPE.itemsize].view(HEADER_DTYPE)[0]
    words = raw[HEADER_DTYPE.itemsize:].view(np.uint64)
    return header, words


def popcount_words(words: np.ndarray, chunk: int = 1 << 26) -> int:
    w = np.asarray(words, dtype=np.uint64).reshape(-1)
    return int(sum(int(np.bitwise_count(w[a:a + chunk]).sum(dtype=np.uint64))
                   for a in range(0, w.size, chunk)))


def load_factor(words: np.ndarray, filter_bits: int) -> float:
    """filter.cuh:810-823: float(popcount of all words) / float(filter bits)."""
    return float(np.float32(popcount_words(words)) / np.float32(filter_bits))
