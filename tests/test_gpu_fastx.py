"""GPU parity of FASTA / FASTQ ingest (insert_fastx / query_fastx families).

Expected values come from the oracles: records from oracle/fastx_oracle.py,
filter words and per-window answers from the CPU Super Bloom oracle on the
records joined by the alphabet separator (a window inside a record gets the
oracle's answer for it; windows never span records), valid windows per
record from the oracle's count. Covers the reference's FASTX tests
(tests/test_fastx.cu, as data), files (plain and gzip), host buffers and
file-like readers, forced chunking (tiny chunks, fill_fraction 0), records
larger than a chunk, parse errors with their location, and a consumer that
stops the query.
"""
from __future__ import annotations

import gzip
import io
import json
import os

import numpy as np
import pytest

from oracle import fastx_oracle as fo
from tests.helpers import make_config, make_oracle, random_dna, random_protein

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_fastx_cases.json")))["cases"]
CONFIGS = {"dna5": (5, 4, 3, 4, "dna"), "triplet3": (3, 2, 2, 4, "triplet")}
SEP = {"dna": b"N", "triplet": b"NNN", "protein": b"*"}


def new_filter(name_or_tuple, bits=1 << 12):
    import cusbf_amd as cs

    k, s, m, h, a = CONFIGS[name_or_tuple] if isinstance(name_or_tuple, str) else name_or_tuple
    return cs.Filter(make_config(k, s, m, h, a), bits)


def write(tmp_path, name, data: bytes, gz=False):
    p = tmp_path / name
    p.write_bytes(gzip.compress(data) if gz else data)
    return str(p)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_reference_fastx_cases(gpu, case, tmp_path):
    from cusbf_amd.errors import CusbfError, ErrorCategory

    text = case["text"].encode()
    path = write(tmp_path, "in.fa", text)
    ff = case.get("fill_fraction", 0.7)
    f = new_filter(case["config"])
    if "error" in case:
        with pytest.raises(CusbfError) as e:
            f.query_fastx_file(path)
        want = case["error"]
        assert e.value.category == ErrorCategory.fastx_parse
        assert (e.value.location.line, e.value.location.column) == (want["line"], want["column"])
        assert e.value.location.file == path
        assert e.value.message == f"{path}:{want['line']}:{want['column']}: {want['detail']}"
        return
    for s in case.get("inserted_before_query") or []:
        f.insert_sequence(s)
    if "insert_report" in case:
        r = f.insert_fastx_file(path, ff)
        assert [r.records_indexed, r.indexed_bases, r.inserted_kmers] == case["insert_report"]
    if "abort_after" in case:
        calls = []

        def consumer(v):
            calls.append((v.record_index, v.header, v.sequence))
            if len(calls) == case["abort_after"]:
                raise RuntimeError("abort")

        with pytest.raises(RuntimeError, match="abort"):
            f.query_fastx_records(io.BytesIO(text), consumer)
        assert calls == [(0, "seqA", b"ACGTACGTACGT")]
        # the filter stays usable
        got = []
        summary = f.query_fastx_records(io.BytesIO(text), lambda v: got.append(v))
        assert [v.record_index for v in got] == [0, 1, 2] and summary.records_queried == 3
        return
    want_q = case.get("query_report") or case.get("query_after_insert_report")
    if want_q is None:
        # tests/test_fastx.cu:33-40: every k-mer of the inserted record answers present
        for _, seq in case["records"]:
            assert f.contains_sequence(seq).all()
        return
    rep = f.query_fastx_file(path, ff)
    assert [rep.records_queried, rep.queried_bases, rep.queried_kmers, rep.positive_kmers] == want_q
    det = f.query_fastx_file_detailed(path, ff)
    assert [(r.header, r.sequence.decode()) for r in det.records] == [tuple(x) for x in case["records"]]
    hits = case.get("record_hits") or case.get("record_hits_after_insert")
    if hits:
        assert [r.hits for r in det.records] == hits
    for r in det.records:
        assert r.positive_kmers == sum(r.hits)
    assert det.summary.queried_kmers == want_q[2]


def fasta_text(records, width=60) -> bytes:
    out = bytearray()
    for i, s in enumerate(records):
        out += b">rec%d some description\n" % i
        for j in range(0, len(s), width):
            out += s[j:j + width] + b"\n"
    return bytes(out)


def fastq_text(records) -> bytes:
    out = bytearray()
    for i, s in enumerate(records):
        out += b"@read%d\r\n" % i + s + b"\r\n+\r\n" + b"I" * len(s) + b"\r\n"
    return bytes(out)


def make_records(alphabet, n, seed, median=300):
    rng = np.random.default_rng(seed)
    lens = np.clip(np.round(rng.lognormal(np.log(median), 1.2, n)), 0, 30000).astype(int)
    lens[:3] = [0, 4, 31]
    gen = random_protein if alphabet == "protein" else random_dna
    kw = {"invalid_rate": 0.002} if alphabet == "protein" else {"n_rate": 0.002}
    return [gen(int(l), seed * 1000 + i, **kw) if l else b"" for i, l in enumerate(lens)]


def expected(cfg_t, records, query_records, num_shards):
    """(words, per-record (valid windows, hits)) from the oracle."""
    k, s, m, h, a = cfg_t
    o = make_oracle(k, s, m, h, a)
    words = o.new_filter(num_shards)
    joined = SEP[a].join(records)
    if joined:
        o.insert(joined, words)
    per = []
    for q in query_records:
        hits = o.contains(q, words) if len(q) >= k else np.zeros(0, np.uint8)
        valid = o.count(q, num_shards)[0] if len(q) >= k else 0
        per.append((valid, hits))
    return words, per


@pytest.mark.parametrize("alphabet,cfg_t,fmt", [
    ("dna", (31, 28, 16, 4, "dna"), "fasta"),
    ("dna", (31, 28, 16, 4, "dna"), "fastq"),
    ("protein", (10, 8, 6, 4, "protein"), "fasta"),
    ("protein", (10, 8, 6, 4, "protein"), "fastq"),
    ("dna", (17, 15, 9, 4, "dna"), "fastq"),  # runtime-span lane strips
])
def test_fastx_file_insert_query_match_oracle(gpu, tmp_path, alphabet, cfg_t, fmt):
    records = make_records(alphabet, 1500, seed=len(fmt) + cfg_t[0])
    gen = random_protein if alphabet == "protein" else random_dna
    # odd records replaced by fresh random symbols of the same length (mostly negatives)
    queries = [r if i % 2 == 0 or not r else gen(len(r), 9000 + i) for i, r in enumerate(records)]
    text = fasta_text(records) if fmt == "fasta" else fastq_text(records)
    qtext = fasta_text(queries) if fmt == "fasta" else fastq_text(queries)
    assert [r.sequence for r in fo.parse(text)] == records
    f = new_filter(cfg_t, bits=max(1 << 16, 16 * len(text)))
    rep = f.insert_fastx_file(write(tmp_path, "in." + fmt, text))
    words, per = expected(cfg_t, records, queries, f.num_shards())
    assert (f.words() == words).all()
    assert rep.records_indexed == len(records)
    assert rep.indexed_bases == sum(len(r) for r in records)
    _, per_ins = expected(cfg_t, records, records, f.num_shards())
    assert rep.inserted_kmers == sum(v for v, _ in per_ins)
    det = f.query_fastx_file_detailed(write(tmp_path, "q." + fmt, qtext))
    assert len(det.records) == len(queries)
    for i, (r, (valid, hits)) in enumerate(zip(det.records, per)):
        assert r.record_index == i and r.sequence == queries[i]
        assert r.queriedKmers == valid, i
        assert np.array_equal(np.array(r.hits, dtype=np.uint8), hits), i
        assert r.positive_kmers == int(hits.sum())
    assert det.summary.positive_kmers == sum(int(h.sum()) for _, h in per)


@pytest.mark.parametrize("fmt,cfg_t", [
    ("fastq", (31, 28, 16, 4, "dna")),
    ("fasta", (31, 28, 16, 4, "dna")),
    ("fasta", (3, 2, 2, 4, "triplet")),  # records padded to whole codons
    ("fastq", (3, 2, 2, 4, "triplet")),
])
def test_fastx_sources_and_chunking_agree(gpu, tmp_path, fmt, cfg_t):
    """File, gzip file, host buffer, reader, tiny chunks, fill_fraction 0: same words and
    answers. FASTA files and buffers are split by host threads, gzip files and readers
    parsed line by line."""
    from cusbf_amd import fastx

    records = make_records("dna", 800, seed=5, median=2000)
    if cfg_t[4] == "triplet":
        records = [r[: len(r) - len(r) % 3 + (1 if i % 5 == 0 else 0)] for i, r in enumerate(records)]
    text = fastq_text(records) if fmt == "fastq" else fasta_text(records, width=70)
    plain = write(tmp_path, "r." + fmt, text)
    gz = write(tmp_path, "r." + fmt + ".gz", text, gz=True)
    results = []
    for mode in ("file", "gzip", "buffer", "reader", "tiny_chunks", "fill0", "record_larger_than_chunk"):
        f = new_filter(cfg_t, bits=1 << 24)
        if mode == "tiny_chunks":
            fastx.set_chunk_bytes(f, 4096)
        if mode == "record_larger_than_chunk":
            fastx.set_chunk_bytes(f, 64)
        ff = 0.0 if mode == "fill0" else 0.7
        if mode == "gzip":
            ri = f.insert_fastx_file(gz, ff)
            rq = f.query_fastx_file_detailed(gz, ff)
        elif mode == "buffer":
            ri = f.insert_fastx(text, ff)
            rq = f.query_fastx_detailed(text, ff)
        elif mode == "reader":
            ri = f.insert_fastx(io.BytesIO(text), ff)
            rq = f.query_fastx_detailed(io.BytesIO(text), ff)
        else:
            ri = f.insert_fastx_file(plain, ff)
            rq = f.query_fastx_file_detailed(plain, ff)
        results.append((mode, f.words(), ri, rq))
    ref_words, ref_ri, ref_rq = results[0][1], results[0][2], results[0][3]
    for mode, words, ri, rq in results[1:]:
        assert (words == ref_words).all(), mode
        assert ri == ref_ri, mode
        assert rq.summary == ref_rq.summary, mode
        assert [(r.header, r.sequence, r.hits) for r in rq.records] == \
               [(r.header, r.sequence, r.hits) for r in ref_rq.records], mode
    assert ref_rq.summary.positive_kmers == ref_rq.summary.queried_kmers  # all inserted


def test_fastx_equals_record_batch_path(gpu):
    """insert_fastx == insert_record_batch of the same records."""
    import cusbf_amd as cs

    cfg = make_config(31, 28, 16, 4)
    records = make_records("dna", 3000, seed=8, median=150)
    a = cs.Filter(cfg, 1 << 26)
    b = cs.Filter(cfg, 1 << 26)
    a.insert_fastx(fasta_text(records, width=70))
    seq = b"".join(records)
    off = np.cumsum([0] + [len(r) for r in records])
    b.insert_record_batch(seq, [(int(off[i]), len(r)) for i, r in enumerate(records)])
    assert (a.words() == b.words()).all()


def test_fastx_parse_error_after_flushed_chunks(gpu, tmp_path):
    """An error after the first chunks: raised with its location; earlier chunks applied."""
    from cusbf_amd import fastx
    from cusbf_amd.errors import CusbfError, ErrorCategory

    records = make_records("dna", 200, seed=3, median=500)
    good = fastq_text(records)
    bad = good + b"@broken\r\nACGT\r\n+\r\nIII\r\n"
    path = write(tmp_path, "bad.fq", bad)
    f = new_filter((31, 28, 16, 4, "dna"), bits=1 << 22)
    fastx.set_chunk_bytes(f, 8192)
    with pytest.raises(CusbfError) as e:
        f.insert_fastx_file(path)
    lines = bad.count(b"\n")
    assert e.value.category == ErrorCategory.fastx_parse
    assert (e.value.location.line, e.value.location.column) == (lines, 4)
    assert "FASTQ quality length does not match sequence length" in e.value.message
    assert f.load_factor() > 0  # chunks before the error were inserted
    # the same filter keeps working
    g = new_filter((31, 28, 16, 4, "dna"), bits=1 << 22)
    g.insert_fastx(good)
    f.insert_fastx(good)
    assert (f.words() == g.words()).all()


def test_fastx_missing_file_is_io_error(gpu, tmp_path):
    from cusbf_amd.errors import CusbfError, ErrorCategory

    f = new_filter("dna5")
    missing = str(tmp_path / "nope.fa")
    with pytest.raises(CusbfError) as e:
        f.insert_fastx_file(missing)
    assert e.value.category == ErrorCategory.io
    assert e.value.message == f"Failed to open FASTA/FASTQ file: {missing}"


@pytest.mark.parametrize("fmt", ["fasta", "fastq"])
def test_consumer_stop_on_split_text(gpu, tmp_path, fmt):
    """A consumer that raises stops a query whose text is split by host threads (a file,
    several chunks): the records before it arrive in order, the call raises, and the
    filter keeps working."""
    from cusbf_amd import fastx

    records = make_records("dna", 600, seed=12, median=400)
    text = fastq_text(records) if fmt == "fastq" else fasta_text(records, width=60)
    path = write(tmp_path, "q." + fmt, text)
    f = new_filter((31, 28, 16, 4, "dna"), bits=1 << 22)
    f.insert_fastx_file(path)
    fastx.set_chunk_bytes(f, 16384)  # many chunks
    seen = []

    def consumer(v):
        seen.append(v.record_index)
        if len(seen) == 250:
            raise RuntimeError("stop here")

    with pytest.raises(RuntimeError, match="stop here"):
        f.query_fastx_file_records(path, consumer)
    assert seen == list(range(250))
    got = []
    rep = f.query_fastx_file_records(path, lambda v: got.append((v.record_index, v.positive_kmers, v.queriedKmers)))
    assert [g[0] for g in got] == list(range(len(records)))
    assert all(p == q for _, p, q in got)  # every record was inserted
    assert rep.records_queried == len(records)
