"""FASTX parsing on the host: the library's native parser against the oracle.

The oracle (oracle/fastx_oracle.py, a line-by-line restatement of the
reference's streaming reader) is pinned by the reference's own FASTX test
inputs (tests/golden/reference_fastx_cases.json, from tests/test_fastx.cu) and
hand-derived cases for every error path; the library's parser must then agree
with it record for record, or error for error (line, column, detail), on those
cases and on seeded random texts with wrapping, CRLF, blank lines and defects.
No device work: cusbf_parse_fastx_buffer only parses.
"""
from __future__ import annotations

import json
import os
import random

import pytest

from oracle import fastx_oracle as fo

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_fastx_cases.json")))["cases"]


def lib_parse(text: bytes):
    """(records, None) or (None, (line, column, detail, message)) from the library."""
    from cusbf_amd import fastx
    from cusbf_amd.errors import CusbfError, ErrorCategory

    try:
        recs = fastx.parse_fastx(text)
        return [(h.encode("latin-1"), s) for h, s in recs], None
    except CusbfError as e:
        assert e.category == ErrorCategory.fastx_parse, e
        detail = e.message.split(": ", 1)[1]
        return None, (e.location.line, e.location.column, detail, e.message, e.location.file)


def oracle_parse(text: bytes):
    recs, err = fo.parse_or_error(text)
    if err is not None:
        return None, (err.line, err.column, err.detail, str(err), err.source)
    return [(r.header, r.sequence) for r in recs], None


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_reference_cases(case):
    text = case["text"].encode()
    o_recs, o_err = oracle_parse(text)
    l_recs, l_err = lib_parse(text)
    if "error" in case:
        want = case["error"]
        assert o_err[:3] == (want["line"], want["column"], want["detail"])
        assert l_err == o_err
        assert l_err[3] == f"<stream>:{want['line']}:{want['column']}: {want['detail']}"
    else:
        want = [(h.encode(), s.encode()) for h, s in case["records"]]
        assert o_recs == want
        assert l_recs == want


# Hand-derived cases for the paths the reference's tests leave open (each
# expectation follows from Fastx.hpp's rules, worked out by hand).
MANUAL = [
    (b"", [], None),
    (b"\n\n\r\n", [], None),
    (b">a\n", [(b"a", b"")], None),                                   # header only
    (b">a", [(b"a", b"")], None),                                      # no newline at all
    (b">a\nAC\n\n>b\r\nGG\r\r\n", [(b"a", b"AC"), (b"b", b"GG\r")], None),  # one '\r' trimmed
    (b"\n\n>x\nA\nC", [(b"x", b"AC")], None),
    (b"@q\nACGT\n+\n@@++\n@r\nA\n+x\nI\n", [(b"q", b"ACGT"), (b"r", b"A")], None),  # '@'/'+' qualities
    (b"@q\n\n+\n\n@r\n\n+\n", [(b"q", b""), (b"r", b"")], None),      # empty FASTQ records
    (b"@q\nAC\nGT\n+\nII\nII\n", [(b"q", b"ACGT")], None),
    (b"ACGT\n", None, (1, 1, "expected FASTA or FASTQ header")),
    (b"\n\n  >a\n", None, (3, 1, "expected FASTA or FASTQ header")),
    (b">a\nAC\n@b\nGG\n", [(b"a", b"AC@bGG")], None),                 # '@' inside FASTA is sequence
    (b"@a\nAC\n+\nII\n>b\nGG\n", None, (5, 1, "mixed FASTA and FASTQ records are not supported")),
    (b"@a\nACGT\n", None, (2, 1, "unterminated FASTQ record: missing '+' separator")),
    (b"@a\nACGT", None, (2, 4, "unterminated FASTQ record: missing '+' separator")),
    (b"@a\nACGT\r\n", None, (2, 1, "unterminated FASTQ record: missing '+' separator")),
    (b"@a\nACGT\nAC", None, (3, 2, "unterminated FASTQ record: missing '+' separator")),
    (b"@a\nACGT\n+\nIIIII\n", None, (4, 5, "FASTQ quality length exceeds sequence length")),
    (b"@a\nACGT\n+\nII\nIIII\n", None, (5, 3, "FASTQ quality length exceeds sequence length")),
    (b"@a\nACGT\n+\nII\n", None, (4, 3, "FASTQ quality length does not match sequence length")),
    (b"@a\nACGT\n+\n", None, (3, 1, "FASTQ quality length does not match sequence length")),
    (b"@a\nACGT\n+\nII\n\n", None, (5, 1, "FASTQ quality length does not match sequence length")),
]


@pytest.mark.parametrize("text,records,error", MANUAL, ids=lambda x: repr(x)[:40])
def test_manual_cases(text, records, error):
    o_recs, o_err = oracle_parse(text)
    l_recs, l_err = lib_parse(text)
    if error is None:
        assert o_err is None and o_recs == records
    else:
        assert o_recs is None and o_err[:3] == error
    assert (l_recs, l_err) == (o_recs, o_err)


def random_text(rng: random.Random) -> bytes:
    fastq = rng.random() < 0.5
    eol = [b"\n", b"\r\n"]
    out = bytearray()
    for r in range(rng.randint(0, 6)):
        for _ in range(rng.choice([0, 0, 0, 1, 2])):
            out += rng.choice(eol)  # blank lines
        seq = bytes(rng.choice(b"ACGTNacgt@+>") for _ in range(rng.randint(0, 40)))
        if not fastq:
            seq = seq.replace(b">", b"A")
        header = (b"@" if fastq else b">") + bytes(rng.choice(b"abc r1:") for _ in range(rng.randint(0, 6)))
        if rng.random() < 0.03:
            header = (b">" if fastq else b"@") + header[1:]  # mixed
        if rng.random() < 0.02:
            header = b"x" + header[1:]
        out += header + rng.choice(eol)
        wrap = rng.choice([1, 3, 7, 80])
        lines = [seq[i:i + wrap] for i in range(0, len(seq), wrap)] or [b""]
        for line in lines:
            if fastq:
                line = line.replace(b"+", b"A")
            out += line + rng.choice(eol)
        if fastq:
            if rng.random() < 0.97:
                out += b"+" + rng.choice([b"", b"r"]) + rng.choice(eol)
            qlen = len(seq)
            if rng.random() < 0.05:
                qlen += rng.choice([-2, -1, 1, 3])
            qual = bytes(rng.choice(b"I#@+5") for _ in range(max(0, qlen)))
            qw = rng.choice([1, 4, 100])
            for i in range(0, len(qual), qw):
                out += qual[i:i + qw] + rng.choice(eol)
    if out and rng.random() < 0.3:
        out = out.rstrip(b"\r\n")
    return bytes(out)


def test_parser_matches_oracle_on_random_texts():
    rng = random.Random(20261203)
    errors = 0
    for i in range(3000):
        text = random_text(rng)
        o = oracle_parse(text)
        l = lib_parse(text)
        assert l == o, (i, text)
        errors += o[1] is not None
    assert 100 < errors < 2000  # both kinds of outcome are exercised


def random_fasta(rng: random.Random, records: int, max_len: int) -> bytes:
    """FASTA with wrapping, CRLF, blank lines, empty and header-only records,
    '@' / '+' / '\\r' inside sequence lines, and sometimes no final newline."""
    eol = [b"\n", b"\r\n"]
    out = bytearray()
    for _ in range(rng.choice([0, 0, 1, 3])):
        out += rng.choice(eol)  # leading blank lines
    for _ in range(records):
        out += b">" + bytes(rng.choice(b"abc r1:>\t") for _ in range(rng.randint(0, 12))) + rng.choice(eol)
        n = rng.choice([0, rng.randint(0, max_len), rng.randint(0, 50)])
        seq = bytes(rng.choice(b"ACGTNacgt@+\r") for _ in range(n))
        wrap = rng.choice([1, 7, 60, 80, 1 << 30])
        for i in range(0, len(seq), wrap):
            out += seq[i:i + wrap] + rng.choice(eol)
            if rng.random() < 0.02:
                out += rng.choice(eol)  # blank line inside the sequence
    if out and rng.random() < 0.3:
        out = out.rstrip(b"\r\n")
    return bytes(out)


@pytest.mark.parametrize("chunk_bytes", [1, 5, 64, 777, 1 << 12, 3 << 20, 0])
@pytest.mark.parametrize("max_records", [0, 1, 3])
def test_fasta_split_in_chunks_matches_oracle(chunk_bytes, max_records):
    """FASTA text in memory is split into records by host threads, chunk by chunk
    (the FASTX pipeline's grouping): the records equal the oracle's for every chunk
    size and record limit, records larger than a chunk included."""
    from cusbf_amd import fastx

    rng = random.Random(7 * chunk_bytes + max_records)
    for i in range(40):
        text = random_fasta(rng, rng.randint(0, 12), 3000)
        want, err = oracle_parse(text)
        assert err is None
        got = fastx.parse_fastx(text, chunk_bytes=chunk_bytes, max_records=max_records)
        assert [(h.encode("latin-1"), s) for h, s in got] == want, (i, text[:200])


def test_fasta_split_large_text_matches_oracle():
    """Texts of several MiB: the split uses parts of about 1 MiB on several host
    threads, with records and lines crossing part boundaries."""
    from cusbf_amd import fastx

    rng = random.Random(99)
    for records, max_len in [(3, 3_000_000), (4000, 3000), (1, 7_000_000)]:
        text = random_fasta(rng, records, max_len)
        want, err = oracle_parse(text)
        assert err is None
        for chunk in (None, 1 << 20, 2_500_001, 10_000):
            got = fastx.parse_fastx(text, chunk_bytes=chunk)
            assert [(h.encode("latin-1"), s) for h, s in got] == want, (records, chunk)


def lib_parse_chunked(text: bytes, chunk_bytes: int, max_records: int = 0):
    from cusbf_amd import fastx
    from cusbf_amd.errors import CusbfError, ErrorCategory

    try:
        recs = fastx.parse_fastx(text, chunk_bytes=chunk_bytes, max_records=max_records)
        return [(h.encode("latin-1"), s) for h, s in recs], None
    except CusbfError as e:
        assert e.category == ErrorCategory.fastx_parse, e
        detail = e.message.split(": ", 1)[1]
        return None, (e.location.line, e.location.column, detail, e.message, e.location.file)


@pytest.mark.parametrize("chunk_bytes", [1, 64, 4096, 0])
@pytest.mark.parametrize("max_records", [0, 2])
def test_chunked_parse_matches_oracle_on_random_texts(chunk_bytes, max_records):
    """FASTQ text in memory is split by host threads where a pattern check and the
    part ends vouch for it, the rest parsed serially: records and errors (line,
    column, detail) equal the oracle's for every chunk size."""
    rng = random.Random(31 * chunk_bytes + max_records + 5)
    for i in range(500):
        text = random_text(rng)
        assert lib_parse_chunked(text, chunk_bytes, max_records) == oracle_parse(text), (i, text)


def random_fastq(rng: random.Random, records: int, max_len: int, wrap: bool = False) -> bytes:
    """Four-line FASTQ (CRLF or LF, blank lines between records, qualities that
    start with '@' or '+'); wrapped sequence and quality lines when `wrap`."""
    eol = rng.choice([b"\n", b"\r\n"])
    out = bytearray()
    for i in range(records):
        if rng.random() < 0.01:
            out += eol  # blank line between records
        n = rng.choice([0, rng.randint(1, max_len), rng.randint(1, 200)])
        seq = bytes(rng.choice(b"ACGTN") for _ in range(n))
        qual = bytes(rng.choice(b"@+I#5") for _ in range(n))
        out += b"@r%d x" % i + eol
        w = rng.choice([7, 60]) if wrap else max(1, n)
        for j in range(0, n, w):
            out += seq[j:j + w] + eol
        if n == 0:
            out += eol
        out += b"+" + eol
        for j in range(0, n, w):
            out += qual[j:j + w] + eol
        if n == 0:
            out += eol
    return bytes(out)


def test_fastq_split_large_texts_match_oracle():
    """Texts of several MiB split into parts of about 1 MiB, with and without
    wrapped lines (which the pattern does not vouch for: parsed serially), and with
    a defect deep inside (reported at its line and column by the serial parser)."""
    rng = random.Random(17)
    for records, max_len, wrap in [(40000, 300, False), (300, 60000, False), (20000, 300, True)]:
        text = random_fastq(rng, records, max_len, wrap)
        want = oracle_parse(text)
        assert want[1] is None
        for chunk in (0, 1 << 20, 300_001):
            assert lib_parse_chunked(text, chunk) == want, (records, wrap, chunk)
    text = random_fastq(rng, 40000, 300)
    lines = text.split(b"\n")
    k = len(lines) * 3 // 4
    while not lines[k].startswith(b"+"):
        k += 1
    lines[k + 1] = lines[k + 1][:-2] + b"\r" if lines[k + 1].endswith(b"\r") else lines[k + 1][:-1]
    bad = b"\n".join(lines)
    want = oracle_parse(bad)
    assert want[1] is not None
    for chunk in (0, 1 << 20, 300_001):
        assert lib_parse_chunked(bad, chunk) == want, chunk
