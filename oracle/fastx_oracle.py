"""FASTA / FASTQ reader restated in plain Python (the oracle for the library's parser).

TEST INFRASTRUCTURE ONLY: the checker the library's native FASTX parser and
FASTX insert/query reports are compared against. Only tests/ import it.

Restates the reference's streaming reader, include/cusbf/Fastx.hpp:
  * lines come from std::getline: split at '\\n', one trailing '\\r' removed
    (trimTrailingCarriageReturn, :254-258); text after the last '\\n' is a line
    when it is not empty; a failing getline at the end of the input leaves the
    previous line in its buffer only when that line had no '\\n';
  * nextRecord (:302-344): empty lines before a header are skipped; the first
    header picks FASTA ('>') or FASTQ ('@'), anything else is "expected FASTA
    or FASTQ header"; a header of the other kind is "mixed FASTA and FASTQ
    records are not supported"; both at column fastx_column_at(header, 0);
  * readFastaSequence (:393-411): lines up to the next line starting with '>';
  * readFastqSequence (:413-434): lines up to a line starting with '+'; the end
    of input first is "unterminated FASTQ record: missing '+' separator" at the
    column of the getline buffer's last byte;
  * readFastqQualities (:436-460): quality lines until their total length
    reaches the sequence length; passing it is "FASTQ quality length exceeds
    sequence length" at fastx_quality_excess_column (:270-281); ending short is
    "FASTQ quality length does not match sequence length" at
    fastx_quality_short_column (:283-287) of the last quality line read.
Errors are (line, column, detail) with 1-based line numbers counting every line
read; the message is "source:line:column: detail" (error.hpp:123-129).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple


class FastxParseError(Exception):
    def __init__(self, source: str, line: int, column: int, detail: str):
        super().__init__(f"{source}:{line}:{column}: {detail}")
        self.source, self.line, self.column, self.detail = source, line, column, detail


def column_at(line: bytes, byte_index: int) -> int:
    """Fastx.hpp:261-268."""
    if not line:
        return 1
    return min(byte_index, len(line) - 1) + 1


def quality_excess_column(quality_length: int, expected_length: int, line: bytes) -> int:
    """Fastx.hpp:270-281."""
    before = quality_length - len(line)
    idx = expected_length - before if expected_length > before else len(line)
    return column_at(line, idx)


def quality_short_column(line: bytes) -> int:
    """Fastx.hpp:283-287."""
    return 1 if not line else len(line) + 1


class _Getline:
    """std::getline over a byte string, with the buffer left by a failing call."""

    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0
        self.buffer = b""  # what the caller's std::string holds after the last call

    def __call__(self) -> Optional[bytes]:
        if self.pos >= len(self.data):
            # sentry fails (eof already set) -> buffer untouched; else erased then nothing read
            if not getattr(self, "_eof_set", False):
                self.buffer = b""
            return None
        nl = self.data.find(b"\n", self.pos)
        if nl < 0:
            line = self.data[self.pos:]
            self.pos = len(self.data)
            self._eof_set = True  # extraction stopped at the end: eofbit
        else:
            line = self.data[self.pos:nl]
            self.pos = nl + 1
        if line.endswith(b"\r"):
            line = line[:-1]
        self.buffer = line
        return line


@dataclass
class Record:
    header: bytes
    sequence: bytes


def parse(data: bytes, source: str = "<stream>") -> List[Record]:
    """Every record of `data`; raises FastxParseError like FastxReader::nextRecord."""
    getline = _Getline(data)
    line_no = 0
    fmt = None
    pending: Optional[bytes] = None
    out: List[Record] = []

    def fail(detail: str, column: int):
        raise FastxParseError(source, line_no, column, detail)

    while True:
        # readHeaderLine
        if pending is not None:
            header, pending = pending, None
        else:
            header = None
            while True:
                line = getline()
                if line is None:
                    break
                line_no += 1
                if line:
                    header = line
                    break
            if header is None:
                return out
        tag = header[:1]
        if fmt is None:
            if tag == b">":
                fmt = "fasta"
            elif tag == b"@":
                fmt = "fastq"
            else:
                fail("expected FASTA or FASTQ header", column_at(header, 0))
        if tag != (b">" if fmt == "fasta" else b"@"):
            fail("mixed FASTA and FASTQ records are not supported", column_at(header, 0))
        seq = bytearray()
        if fmt == "fasta":
            while True:
                line = getline()
                if line is None:
                    break
                line_no += 1
                if line[:1] == b">":
                    pending = line
                    break
                seq += line
        else:
            while True:
                line = getline()
                if line is None:
                    buf = getline.buffer
                    fail("unterminated FASTQ record: missing '+' separator",
                         column_at(buf, len(buf) - 1 if len(buf) > 0 else 0))
                line_no += 1
                if line[:1] == b"+":
                    break
                seq += line
            expected = len(seq)
            got = 0
            short_column = 1
            while got < expected:
                line = getline()
                if line is None:
                    break
                line_no += 1
                short_column = quality_short_column(line)
                got += len(line)
                if got > expected:
                    fail("FASTQ quality length exceeds sequence length",
                         quality_excess_column(got, expected, line))
            if got != expected:
                fail("FASTQ quality length does not match sequence length", short_column)
        out.append(Record(header[1:], bytes(seq)))


def record_windows(seq: bytes, k: int, width: int = 1) -> int:
    """record_kmer_count (record_math.cuh:12-16): windows of one record."""
    symbols = len(seq) // width
    return symbols - k + 1 if symbols >= k else 0


def parse_or_error(data: bytes, source: str = "<stream>") -> Tuple[Optional[List[Record]],
                                                                   Optional[FastxParseError]]:
    try:
        return parse(data, source), None
    except FastxParseError as e:
        return None, e
