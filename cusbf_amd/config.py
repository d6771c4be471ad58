"""Config and alphabets (reference: include/cusbf/config.cuh, include/cusbf/Alphabet.cuh).

`Config(k, s, m, hash_count=4, alphabet=DnaAlphabet)` is the runtime form of
cusbf::Config<K, S, M, HashCount, CudaBlockSize, Alphabet>. CudaBlockSize has
no meaning for the wave64 kernels and is accepted and ignored.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Optional

from . import _lib
from .errors import check

ALPHABET_DNA = 0
ALPHABET_DNA_TRIPLET = 1
ALPHABET_PROTEIN = 2
ALPHABET_CUSTOM = 3


@dataclass(frozen=True)
class Alphabet:
    """An alphabet (Alphabet.cuh:94-107).

    Built-in alphabets carry their library code; a custom width-1 alphabet is
    defined by `encode(byte) -> symbol` (evaluated once for all 256 bytes),
    its symbol count, invalid sentinel and separator byte.
    """

    name: str
    code: int
    symbol_width: int
    symbol_count: int
    invalid_symbol: int
    separator: int
    valid_bytes: bytes
    encode: Optional[Callable[[int], int]] = None

    def lut(self) -> bytes:
        if self.encode is None:
            return bytes(256)
        return bytes(self.encode(b) & 0xFF for b in range(256))

    @staticmethod
    def custom(name: str, symbol_count: int, invalid_symbol: int, separator: str | int,
               valid_bytes: bytes, encode: Callable[[int], int]) -> "Alphabet":
        sep = ord(separator) if isinstance(separator, str) else int(separator)
        alpha = Alphabet(name, ALPHABET_CUSTOM, 1, symbol_count, invalid_symbol, sep,
                         bytes(valid_bytes), encode)
        # Alphabet.cuh:186-195: the separator must never encode as a valid symbol.
        if encode(sep) != invalid_symbol:
            raise ValueError("the separator byte must encode as the invalid symbol")
        return alpha


# Alphabet.cuh:113-142
DnaAlphabet = Alphabet("DnaAlphabet", ALPHABET_DNA, 1, 4, 0xFF, ord("N"), b"ACGT")
# Alphabet.cuh:150-178
DnaTripletAlphabet = Alphabet("DnaTripletAlphabet", ALPHABET_DNA_TRIPLET, 3, 64, 0xFF, ord("N"),
                              b"ACGT")
# Alphabet.cuh:189-218
ProteinAlphabet = Alphabet("ProteinAlphabet", ALPHABET_PROTEIN, 1, 26, 0xFF, ord("*"),
                           b"ABCDEFGHIJKLMNOPQRSTUVWXYZ")


class Config:
    """config.cuh:28-118; invalid parameters raise CusbfError(invalid_argument)."""

    def __init__(self, k: int, s: int, m: int, hash_count: int = 4, cuda_block_size: int = 256,
                 alphabet: Alphabet = DnaAlphabet):
        self.k, self.s, self.m, self.hash_count = int(k), int(s), int(m), int(hash_count)
        self.alphabet = alphabet
        self.cuda_block_size = cuda_block_size
        lib = _lib.load()
        c = _lib.CConfig()
        if alphabet.code == ALPHABET_CUSTOM:
            c.k, c.s, c.m, c.hash_count = self.k, self.s, self.m, self.hash_count
            c.alphabet = ALPHABET_CUSTOM
            c.symbol_count = alphabet.symbol_count
            c.invalid_symbol = alphabet.invalid_symbol
            c.separator = alphabet.separator
            ctypes.memmove(c.lut, alphabet.lut(), 256)
        else:
            check(lib, lib.cusbf_config_init(ctypes.byref(c), self.k, self.s, self.m,
                                             self.hash_count, alphabet.code))
        self._c = c
        bits = ctypes.c_uint32()
        check(lib, lib.cusbf_config_symbol_bits(ctypes.byref(c), ctypes.byref(bits)))
        self.symbol_bits = bits.value
        self.symbol_width = alphabet.symbol_width
        self.symbol_mask = (1 << self.symbol_bits) - 1
        self.minimizer_span = self.k - self.m + 1
        self.findere_span = self.k - self.s + 1
        self.filter_block_bits = 256
        self.block_word_count = 4

    @property
    def c_config(self) -> _lib.CConfig:
        return self._c

    @staticmethod
    def from_c(c: _lib.CConfig) -> "Config":
        """The Config a cusbf_config describes (filter images store one)."""
        by_code = {ALPHABET_DNA: DnaAlphabet, ALPHABET_DNA_TRIPLET: DnaTripletAlphabet,
                   ALPHABET_PROTEIN: ProteinAlphabet}
        if c.alphabet in by_code:
            alphabet = by_code[c.alphabet]
        else:
            lut = bytes(c.lut)
            alphabet = Alphabet.custom("image", int(c.symbol_count), int(c.invalid_symbol),
                                       int(c.separator), b"", lambda b: lut[b])
        return Config(int(c.k), int(c.s), int(c.m), int(c.hash_count), alphabet=alphabet)

    def __repr__(self) -> str:
        return (f"Config<{self.k},{self.s},{self.m},{self.hash_count},{self.alphabet.name}>")
