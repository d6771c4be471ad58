"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact filter words and hit bytes on seeded inputs, the reference's own
known-answer tests (tests/test_bloom_filter.cu), the dense packed path, record
batches, unaligned buffers, tile seams and the fast-vs-generic kernel split.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests.helpers import (KNOWN, config_for, make_config, make_oracle, oracle_for,
                           random_bytes_for, random_dna)

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def dev(data: bytes, offset: int = 0):
    """Device copy of `data`, optionally starting `offset` bytes into the allocation."""
    torch = _torch()
    buf = torch.zeros(len(data) + offset + 16, dtype=torch.uint8, device="cuda")
    if data:
        buf[offset:offset + len(data)] = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    return buf[offset:offset + len(data)]


def new_filter(cfg, bits):
    import cusbf_amd as cs

    return cs.Filter(cfg, bits)


# --- the reference's known-answer tests, through the product ------------------------------------

@pytest.mark.parametrize("case", KNOWN["insert_then_query"], ids=lambda c: c["source"])
def test_known_answers(gpu, case):
    f = new_filter(config_for(case["config"]), case["filter_bits"])
    assert f.insert_sequence(case["insert"]) == case["inserted"]
    hits = f.contains_sequence(case["query"])
    if case["hits"] == "all_ones":
        assert hits.size == case["inserted"] and hits.all()
    else:
        assert hits.tolist() == case["hits"]
    o = oracle_for(case["config"])
    w = o.new_filter(f.num_shards())
    o.insert(case["insert"], w)
    assert (f.words() == w).all()


def test_device_record_batch_known_answer(gpu):
    torch = _torch()
    c = KNOWN["device_record_batch"]
    cfg = config_for(c["config"])
    batch = (c["record1"] + c["separator"] + c["record2"]).encode()
    f = new_filter(cfg, c["filter_bits"])
    assert f.insert_record_batch_async(dev(batch), c["ranges"]) == c["inserted"]
    torch.cuda.synchronize()
    assert f.contains_sequence(c["record1"]).tolist() == [1] * c["record_hits"]
    assert f.contains_sequence(c["record2"]).tolist() == [1] * c["record_hits"]
    g = new_filter(cfg, c["filter_bits"])
    g.insert_sequence(c["record1"])
    g.insert_sequence(c["record2"])
    out = torch.full((len(c["batch_hits"]),), 7, dtype=torch.uint8, device="cuda")
    g.contains_record_batch_async(dev(batch), c["ranges"], out)
    torch.cuda.synchronize()
    assert out.cpu().tolist() == c["batch_hits"]


@pytest.mark.parametrize("case", KNOWN["dense_equals_bytes"], ids=lambda c: c["source"])
def test_dense_equals_bytes_known(gpu, case):
    import cusbf_amd as cs

    cfg = config_for(case["config"])
    f = new_filter(cfg, case["filter_bits"])
    seq = case["sequence"]
    packed = cs.pack_dense_sequence(cfg, seq)
    assert f.insert_dense_packed(packed, len(seq)) == case["inserted"]
    hits = f.contains_dense_packed(packed, len(seq))
    assert hits.size == case["inserted"] and hits.all()
    assert (f.contains_sequence(seq) == hits).all()


def test_short_sequence_leaves_device_output_untouched(gpu):
    torch = _torch()
    f = new_filter(config_for("TestConfig"), 1 << 12)
    out = torch.full((1,), 0xAB, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(dev(b"ACGT"), out)
    torch.cuda.synchronize()
    assert out.item() == 0xAB
    assert f.insert_sequence("ACGT") == 0
    assert f.contains_sequence("ACGT").size == 0


def test_clear_and_idempotence(gpu):
    f = new_filter(config_for("TestConfig"), 1 << 12)
    seq = "ACGTACGTACGTACGT"
    f.insert_sequence(seq)
    lf1 = f.load_factor()
    f.insert_sequence(seq)
    assert f.load_factor() == lf1 > 0
    f.clear()
    assert f.load_factor() == 0.0
    assert not f.contains_sequence(seq).any()


@pytest.mark.parametrize("h", [4, 16])
def test_load_factor_equals_oracle_popcount(gpu, h):
    """load_factor (filter.cuh:810-823) == float(popcount of the oracle's words) /
    float(filter bits): sliced (h = 4) and salted (h = 16) bit addresses."""
    from oracle import image_oracle as im

    cfg = make_config(31, 28, 16, h)
    o = make_oracle(31, 28, 16, h)
    seq = random_dna(700_000, seed=40 + h, n_rate=0.001)
    f = new_filter(cfg, 6 * len(seq))
    f.insert_sequence(seq)
    w = o.new_filter(f.num_shards())
    o.mt_insert(seq, w)
    want = im.load_factor(w, f.filter_bits())
    assert 0.0 < want < 1.0
    assert f.load_factor() == want


def test_host_and_device_query_agree(gpu):
    torch = _torch()
    c = KNOWN["host_vs_device"]
    f = new_filter(config_for(c["config"]), c["filter_bits"])
    f.insert_sequence(c["insert"])
    host = f.contains_sequence(c["query"])
    out = torch.empty(host.size, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(dev(c["query"].encode()), out)
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == host).all()


# --- seeded random parity, bit-exact words and answers ------------------------------------------

FAST = [("dna", 31, 28, 16, 4), ("dna", 31, 30, 16, 4), ("dna", 31, 31, 16, 4),
        ("dna", 31, 28, 21, 4), ("dna", 31, 28, 16, 8), ("dna", 31, 28, 16, 12),
        ("dna", 31, 28, 16, 16), ("dna", 31, 28, 24, 4), ("dna", 31, 30, 20, 8),
        ("dna", 31, 31, 31, 4), ("dna", 5, 4, 3, 4), ("dna", 3, 2, 2, 4),
        ("dna", 21, 19, 11, 4), ("protein", 10, 8, 6, 4), ("protein", 12, 10, 6, 4),
        ("protein", 5, 4, 3, 4)]
GENERIC = [("dna", 27, 25, 13, 4), ("dna", 32, 20, 32, 8), ("dna", 31, 24, 16, 12),
           ("dna", 31, 28, 18, 16), ("dna", 32, 16, 8, 4), ("protein", 12, 8, 5, 16),
           ("triplet", 7, 5, 4, 4), ("triplet", 10, 9, 3, 12), ("triplet", 10, 6, 2, 16),
           ("custom_xyz", 12, 10, 6, 4), ("custom_xyz", 20, 14, 9, 8),
           ("custom_ab", 64, 40, 20, 8), ("custom_ab", 64, 64, 64, 4), ("custom_ab", 64, 33, 1, 16)]


def _parity(alphabet, k, s, m, h, n, seed, invalid_rate, bits_per_kmer=8.0, offset=0):
    torch = _torch()
    cfg = make_config(k, s, m, h, alphabet)
    o = make_oracle(k, s, m, h, alphabet)
    seq = random_bytes_for(alphabet, n, seed, invalid_rate)
    if alphabet in ("dna", "triplet"):
        b = bytearray(seq)
        rng = np.random.default_rng(seed + 1)
        for i in rng.integers(0, n, n // 50):
            b[i] |= 0x20  # lower case
        seq = bytes(b)
    f = new_filter(cfg, int(bits_per_kmer * max(n, 1)))
    f.insert_sequence_async(dev(seq, offset))
    torch.cuda.synchronize()
    w = o.new_filter(f.num_shards())
    o.mt_insert(seq, w)
    got = f.words()
    assert got.shape == w.shape
    bad = np.nonzero(got != w)[0]
    assert bad.size == 0, f"{bad.size} filter words differ (first shard {bad[0] // 4})"
    q = random_bytes_for(alphabet, n // 2, seed + 7, invalid_rate) + seq[: n // 2]
    out = torch.empty(max(1, o.kmer_count(len(q))), dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(dev(q, offset), out)
    torch.cuda.synchronize()
    want = o.mt_contains(q, w)
    got_hits = out.cpu().numpy()[: want.size]
    bad = np.nonzero(got_hits != want)[0]
    assert bad.size == 0, f"{bad.size} answers differ (first window {bad[0]})"
    return f


@pytest.mark.parametrize("alphabet,k,s,m,h", FAST)
def test_lane_strip_parity(gpu, alphabet, k, s, m, h):
    f = _parity(alphabet, k, s, m, h, 300_000, seed=k * 131 + s * 7 + m + h, invalid_rate=0.003)
    assert f.kernel_family() == "lane_strip"


@pytest.mark.parametrize("form", ["direct", "hashed"])
@pytest.mark.parametrize("alphabet,k,s,m,h", GENERIC)
def test_generic_parity(gpu, monkeypatch, form, alphabet, k, s, m, h):
    # both forms of the generic kernels (per-window recomputation, hashed tiles
    # with sparse tables) against the oracle, whichever the spans would pick;
    # configurations the runtime-span lane strips cover are forced onto them
    monkeypatch.setenv("CUSBF_GENERIC_FORM", form)
    monkeypatch.setenv("CUSBF_FORCE_GENERIC", "1")
    f = _parity(alphabet, k, s, m, h, 60_000, seed=k * 17 + h, invalid_rate=0.003)
    assert f.kernel_family() == "generic"


@pytest.mark.parametrize("offset", [1, 3, 8, 13])
@pytest.mark.parametrize("alphabet,k,s,m", [("dna", 31, 28, 16), ("protein", 10, 8, 6)])
def test_unaligned_device_buffers(gpu, offset, alphabet, k, s, m):
    _parity(alphabet, k, s, m, 4, 50_001, seed=offset, invalid_rate=0.002, offset=offset)


@pytest.mark.parametrize("n", [31, 32, 46, 47, 991 + 30, 992 + 30, 993 + 30, 3968 + 30,
                               3969 + 30, 4096, 65_535, 65_536 + 29])
def test_tile_seams(gpu, n):
    _parity("dna", 31, 28, 16, 4, n, seed=n, invalid_rate=0.0, bits_per_kmer=64)


@pytest.mark.parametrize("n", [10, 11, 25, 26, 1008 + 9, 1009 + 9, 4032 + 9, 4096, 65_536 + 9])
def test_tile_seams_protein(gpu, n):
    # protein k=10: one halo lane, 63 window lanes, 1008 windows per wave tile
    _parity("protein", 10, 8, 6, 4, n, seed=n, invalid_rate=0.0, bits_per_kmer=64)


def test_invalid_heavy_input(gpu):
    _parity("dna", 31, 28, 16, 4, 200_000, seed=9, invalid_rate=0.05)


def test_all_invalid_input(gpu):
    torch = _torch()
    f = new_filter(make_config(31, 28, 16, 4), 1 << 16)
    seq = b"N" * 5000
    assert f.insert_sequence_async(dev(seq)) == 5000 - 30
    torch.cuda.synchronize()
    assert not f.words().any()
    assert not f.contains_sequence(seq).any()


def test_fast_matches_generic_kernels(gpu, monkeypatch):
    import cusbf_amd as cs

    torch = _torch()
    seq = random_dna(400_000, seed=77, n_rate=0.001)
    cfg = make_config(31, 28, 16, 4)
    fast = cs.Filter(cfg, 1 << 23)
    monkeypatch.setenv("CUSBF_FORCE_GENERIC", "1")
    slow = cs.Filter(cfg, 1 << 23)
    assert fast.kernel_family() == "lane_strip" and slow.kernel_family() == "generic"
    d = dev(seq)
    fast.insert_sequence_async(d)
    slow.insert_sequence_async(d)
    torch.cuda.synchronize()
    assert (fast.words() == slow.words()).all()


# --- dense packed path ---------------------------------------------------------------------------

@pytest.mark.parametrize("alphabet,k,s,m,h", [("dna", 31, 28, 16, 4), ("dna", 5, 4, 3, 4),
                                              ("protein", 10, 8, 6, 4), ("triplet", 7, 5, 4, 4),
                                              ("custom_xyz", 12, 10, 6, 4)])
def test_dense_parity(gpu, alphabet, k, s, m, h):
    import cusbf_amd as cs

    torch = _torch()
    cfg = make_config(k, s, m, h, alphabet)
    o = make_oracle(k, s, m, h, alphabet)
    seq = random_bytes_for(alphabet, 120_003, seed=k + h, invalid_rate=0.002)
    nsym = len(seq) // cfg.symbol_width
    host_packed = o.pack_dense(seq)
    assert (cs.pack_dense_sequence(cfg, seq) == host_packed).all()
    d_words = torch.zeros(host_packed.size, dtype=torch.int64, device="cuda")
    cs.pack_dense_sequence_device(cfg, dev(seq), nsym, d_words)
    torch.cuda.synchronize()
    assert (d_words.cpu().numpy().view(np.uint64) == host_packed).all()
    f = cs.Filter(cfg, 8 * nsym)
    f.insert_dense_packed_async(d_words, nsym)
    torch.cuda.synchronize()
    w = o.new_filter(f.num_shards())
    o.insert_dense(host_packed, nsym, w)
    assert (f.words() == w).all()
    out = torch.empty(nsym - k + 1, dtype=torch.uint8, device="cuda")
    f.contains_dense_packed_async(d_words, nsym, out)
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == o.contains_dense(host_packed, nsym, w)).all()


def test_dense_span_too_small_is_invalid_argument(gpu):
    import cusbf_amd as cs

    torch = _torch()
    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 16)
    d = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(cs.CusbfError) as e:
        f.insert_dense_packed_async(d, 100)
    assert e.value.category == cs.ErrorCategory.invalid_argument
    assert e.value.message == "dense packed span is too small for num_symbols"


# --- host record batches ------------------------------------------------------------------------

def test_host_record_batch_reports(gpu):
    import cusbf_amd as cs

    cfg = make_config(31, 28, 16, 4)
    o = make_oracle(31, 28, 16, 4)
    rng = np.random.default_rng(4)
    recs = [random_dna(int(x), seed=i, n_rate=0.002) for i, x in enumerate(rng.integers(10, 3000, 300))]
    seq = b"".join(recs)
    ranges, off = [], 0
    for r in recs:
        ranges.append((off, len(r)))
        off += len(r)
    f = cs.Filter(cfg, 1 << 22)
    rep = f.insert_record_batch(seq, ranges)
    norm, meta = cs.normalize_record_batch(cfg, seq, ranges)
    assert rep.records_indexed == len(recs)
    assert rep.indexed_bases == len(seq)
    assert rep.inserted_kmers == sum(r.valid_kmers for r in meta)
    w = o.new_filter(f.num_shards())
    o.insert(norm, w)
    assert (f.words() == w).all()
    q = f.query_record_batch(seq, ranges)
    assert q.queried_kmers == rep.inserted_kmers == q.positive_kmers
    with pytest.raises(cs.CusbfError) as e:
        f.insert_record_batch(seq, [(10, 5), (0, 5)])
    assert e.value.category == cs.ErrorCategory.invalid_argument


def test_count_windows_and_runs_matches_oracle(gpu):
    torch = _torch()
    o = make_oracle(31, 28, 16, 4)
    seq = random_dna(500_000, seed=31, n_rate=0.001)
    f = new_filter(make_config(31, 28, 16, 4), 1 << 24)
    assert f.count_windows_runs(dev(seq)) == o.count(seq, f.num_shards())
    torch.cuda.synchronize()


# --- full-size properties ------------------------------------------------------------------------

def test_large_insert_query_no_false_negatives_and_oracle_words(gpu):
    import cusbf_amd as cs

    torch = _torch()
    n = 64_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    codes = torch.randint(0, 4, (n,), dtype=torch.uint8, device="cuda", generator=g)
    lut = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device="cuda")
    seq = lut[codes.long()]
    cfg = make_config(31, 28, 16, 4)
    f = cs.Filter(cfg, int(21.4 * n))
    f.insert_sequence_async(seq)
    out = torch.empty(n - 30, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(seq, out)
    torch.cuda.synchronize()
    assert int(out.sum().item()) == n - 30
    host = seq.cpu().numpy().tobytes()
    o = make_oracle(31, 28, 16, 4)
    w = o.new_filter(f.num_shards())
    o.mt_insert(host, w)
    assert (f.words() == w).all()


@pytest.mark.parametrize("s", [28, 30, 31])
def test_fpr_on_seeded_negatives_equals_oracle(gpu, s):
    """The small plain workload's FPR protocol: insert one seeded uniform 10 Mbp
    sequence (filter requested at 21.4 bits per k-mer), query it (every k-mer
    present) and an independent seeded 10 Mbp sequence; the GPU's positives
    on the negatives equal the oracle's exactly, so the empirical FPR matches."""
    import cusbf_amd as cs

    torch = _torch()
    n = 10_000_000
    a = random_dna(n, seed=101)
    b = random_dna(n, seed=202)
    cfg = make_config(31, s, 16, 4)
    f = cs.Filter(cfg, int(21.4 * (n - 30)))
    f.insert_sequence(a)
    hits_a = f.contains_sequence(a)
    assert int(np.count_nonzero(hits_a)) == n - 30
    hits_b = f.contains_sequence(b)
    o = make_oracle(31, s, 16, 4)
    w = o.new_filter(f.num_shards())
    o.mt_insert(a, w)
    want = o.mt_contains(b, w)
    assert (np.asarray(hits_b, dtype=np.uint8) == want.astype(np.uint8)).all()
    fpr = np.count_nonzero(want) / (n - 30)
    assert 0 < fpr < 0.02, fpr
    torch.cuda.synchronize()


def test_full_size_query_answers_equal_oracle(gpu):
    """Query answers at full size: the 1M-record DNA workload joined by separators
    (~3.08 G windows) is inserted, then the same records with every odd record replaced
    by fresh symbols are queried; all ~3.08 G answer bytes equal the multithreaded
    oracle's, positives and negatives both present."""
    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    torch = _torch()
    torch.cuda.empty_cache()
    spec = W.BatchedSpec(records=1_000_000, letters=W.DNA_LETTERS, seed=31)
    data, _, lens = W.generate_batched(spec)
    query = W.mutate_odd_records(data, lens, W.DNA_LETTERS, spec.seed + 1)
    text = W.join_records(data, lens, ord("N"))
    qtext = W.join_records(query, lens, ord("N"))
    del data, query
    cfg = make_config(31, 28, 16, 4)
    f = cs.Filter(cfg, int(21.4 * text.numel()))
    f.insert_sequence_async(text)
    out = torch.empty(qtext.numel() - 30, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(qtext, out)
    got = out.cpu().numpy()
    host_t, host_q = text.cpu().numpy(), qtext.cpu().numpy()
    del out, text, qtext
    torch.cuda.synchronize()
    o = make_oracle(31, 28, 16, 4)
    w = o.new_filter(f.num_shards())
    o.mt_insert(host_t, w)
    want = o.mt_contains(host_q, w)
    assert np.array_equal(got, want)
    positives = int(np.count_nonzero(want))
    assert 0 < positives < want.size
