"""Pins the CPU oracle (CPU only).

The oracle is checked against (a) every known-answer test of the reference
that covers the path (tests/golden/reference_known_answers.json), (b)
published constants, (c) brute force on small inputs written from the
definitions (string slicing and base-4 digits, not the oracle's packed shifts),
and (d) invariants (no false negatives, order independence, idempotence,
dense == byte path, naive == fast form).
"""
from __future__ import annotations

import itertools
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import (BASE_CODE, KNOWN, make_oracle, oracle_for, random_bytes_for,
                           random_dna)

MASK64 = (1 << 64) - 1


# --- (a) the reference's own known answers ----------------------------------------------------

@pytest.mark.parametrize("case", KNOWN["insert_then_query"], ids=lambda c: c["source"])
def test_known_insert_then_query(case):
    o = oracle_for(case["config"])
    w = o.new_filter(o.num_shards(case["filter_bits"]))
    o.insert(case["insert"], w)
    assert o.kmer_count(len(case["insert"])) == case["inserted"]
    hits = o.contains(case["query"], w)
    if case["hits"] == "all_ones":
        assert hits.size == case["inserted"] and hits.all()
    else:
        assert hits.tolist() == case["hits"]


def test_known_device_record_batch():
    c = KNOWN["device_record_batch"]
    o = oracle_for(c["config"])
    w = o.new_filter(o.num_shards(c["filter_bits"]))
    batch = c["record1"] + c["separator"] + c["record2"]
    o.insert(batch, w)
    assert sum(o.kmer_count(r[1]) for r in c["ranges"]) == c["inserted"]
    assert o.contains(c["record1"], w).tolist() == [1] * c["record_hits"]
    assert o.contains(c["record2"], w).tolist() == [1] * c["record_hits"]
    w2 = o.new_filter(o.num_shards(c["filter_bits"]))
    o.insert(c["record1"], w2)
    o.insert(c["record2"], w2)
    assert o.contains(batch, w2).tolist() == c["batch_hits"]


@pytest.mark.parametrize("case", KNOWN["dense_equals_bytes"], ids=lambda c: c["source"])
def test_known_dense_equals_bytes(case):
    o = oracle_for(case["config"])
    n = o.num_shards(case["filter_bits"])
    seq = case["sequence"]
    packed = o.pack_dense(seq)
    wd = o.new_filter(n)
    o.insert_dense(packed, len(seq), wd)
    hits = o.contains_dense(packed, len(seq), wd)
    assert hits.size == case["inserted"] and hits.all()
    assert (o.contains(seq, wd) == hits).all()
    wb = o.new_filter(n)
    o.insert(seq, wb)
    assert (wb == wd).all()


def test_known_multiple_inserts_and_idempotence():
    c = KNOWN["multiple_inserts"]
    o = oracle_for(c["config"])
    w = o.new_filter(o.num_shards(c["filter_bits"]))
    for s in c["sequences"]:
        o.insert(s, w)
    for s in c["sequences"]:
        assert o.contains(s, w).all()
    before = w.copy()
    o.insert(c["sequences"][0], w)
    assert (before == w).all()  # RepeatedInsertionIsIdempotent (test_bloom_filter.cu:103-115)


# --- (b) constants ----------------------------------------------------------------------------

def test_golden_ratio_multiplier():
    getcontext().prec = 60
    phi = (Decimal(1) + Decimal(5).sqrt()) / 2
    # floor(2^64 / phi): the odd integer just below 2^64 / phi (Knuth's multiplier).
    from decimal import ROUND_FLOOR
    assert int((Decimal(2) ** 64 / phi).to_integral_value(rounding=ROUND_FLOOR)) == 0x9E3779B97F4A7C15
    assert orc.minimizer_hash64(1) == 0x9E3779B97F4A7C15
    assert orc.salt(0) == 0x9E3779B97F4A7C15


def test_salts_match_published_constants():
    # xxHash64 PRIME64_2..PRIME64_5 (xxHash specification) and the first
    # splitmix64 multiplier; salt 5 is transcribed from the reference as is.
    assert [orc.salt(i) for i in (1, 2, 3, 4)] == [
        0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9, 0x85EBCA77C2B2AE63, 0x27D4EB2F165667C5]
    assert orc.salt(6) == 0xBF58476D1CE4E5B9
    assert len({orc.salt(i) for i in range(16)}) == 16
    assert all(orc.salt(i) & 1 for i in range(16))  # odd multipliers


def test_hash64_definition():
    rng = np.random.default_rng(7)
    for x in [0, 1, 2, MASK64, *map(int, rng.integers(0, 2**63, 50))]:
        p = (x * 0x9E3779B97F4A7C15) & MASK64
        assert orc.hash64(x) == p ^ (p >> 33)
        assert orc.minimizer_hash64(x) == p


# --- (c) brute force from the definitions -------------------------------------------------------

def _packed(s: str) -> int:
    v = 0
    for ch in s:
        v = v * 4 + BASE_CODE[ch.upper()]
    return v


def _brute_window(s: str, k: int, sm: int, m: int, h: int):
    """Shard hash and masks of one DNA window from string slices."""
    mins = min((_packed(s[o:o + m]) * 0x9E3779B97F4A7C15) & MASK64 for o in range(k - m + 1))
    masks = [0, 0, 0, 0]
    slice_w = 64 // h
    for o in range(k - sm + 1):
        p = (_packed(s[o:o + sm]) * 0x9E3779B97F4A7C15) & MASK64
        hv = p ^ (p >> 33)
        for i in range(h):
            if slice_w >= 6:
                pos = (hv >> (slice_w * i)) & 63
            else:
                pos = ((hv * orc.salt(i)) & MASK64) >> 58
            masks[i % 4] |= 1 << pos
    return mins, masks


@pytest.mark.parametrize("k,s,m,h", [(5, 4, 3, 4), (31, 28, 16, 4), (31, 28, 21, 8),
                                     (21, 19, 11, 12), (31, 31, 16, 16), (32, 20, 32, 4)])
def test_single_kmer_bits_lie_in_its_minimizer_shard(k, s, m, h):
    o = make_oracle(k, s, m, h)
    shards = 1 << 10
    rng = np.random.default_rng(k * 100 + h)
    for _ in range(20):
        kmer = "".join(rng.choice(list("ACGT"), k))
        mh, masks = _brute_window(kmer, k, s, m, h)
        w = o.new_filter(shards)
        o.insert(kmer, w)
        shard = mh & (shards - 1)
        nz = np.nonzero(w)[0]
        assert set((nz // 4).tolist()) <= {shard}
        assert w[shard * 4:shard * 4 + 4].tolist() == masks
        valid, packed, omh, omasks = o.window(kmer, 0)
        assert valid and packed == _packed(kmer) and omh == mh and omasks.tolist() == masks


@pytest.mark.parametrize("k,s,m,h", [(64, 40, 20, 8), (64, 64, 64, 4), (64, 33, 1, 16)])
def test_one_bit_alphabet_up_to_k64(k, s, m, h):
    # a two-symbol custom alphabet packs one bit per symbol, so a k-mer fills the
    # whole uint64_t at k = 64 (window_mask at 64 bits, shifts by 0 and 63)
    from helpers import ab_encode

    o = make_oracle(k, s, m, h, "custom_ab")
    rng = np.random.default_rng(k + s + m)

    def packed(t: str) -> int:
        v = 0
        for ch in t:
            v = (v << 1) | ab_encode(ord(ch))
        return v

    for _ in range(20):
        kmer = "".join(rng.choice(list("ab"), k))
        mh = min((packed(kmer[i:i + m]) * 0x9E3779B97F4A7C15) & MASK64 for i in range(k - m + 1))
        masks = [0, 0, 0, 0]
        for i in range(k - s + 1):
            p = (packed(kmer[i:i + s]) * 0x9E3779B97F4A7C15) & MASK64
            hv = p ^ (p >> 33)
            for j in range(h):
                sw = 64 // h
                pos = (hv >> (sw * j)) & 63 if sw >= 6 else ((hv * orc.salt(j)) & MASK64) >> 58
                masks[j % 4] |= 1 << pos
        valid, pk, omh, omasks = o.window(kmer, 0)
        assert valid and pk == packed(kmer) and omh == mh and omasks.tolist() == masks


def test_exhaustive_small_kmers_have_no_false_negatives():
    o = make_oracle(5, 4, 3, 4)
    shards = o.num_shards(1 << 12)
    for kmer in itertools.product("ACGT", repeat=5):
        s = "".join(kmer)
        w = o.new_filter(shards)
        o.insert(s, w)
        assert o.contains(s, w).tolist() == [1]
        mh, masks = _brute_window(s, 5, 4, 3, 4)
        sh = mh & (shards - 1)
        assert w[sh * 4:sh * 4 + 4].tolist() == masks


def test_bit_address_slices_and_salts():
    rng = np.random.default_rng(3)
    for h in (4, 8, 12, 16):
        o = make_oracle(31, 28, 16, h)
        for hv in map(int, rng.integers(0, 2**63, 20, dtype=np.int64)):
            hv = (hv << 1) | 1
            for i in range(h):
                sw = 64 // h
                want = (hv >> (sw * i)) & 63 if sw >= 6 else ((hv * orc.salt(i)) & MASK64) >> 58
                assert o.bit_address(i, hv) == want


def test_num_shards_power_of_two_rounding():
    assert orc.Oracle.num_shards(0) == 1
    assert orc.Oracle.num_shards(1) == 1
    assert orc.Oracle.num_shards(256) == 1
    assert orc.Oracle.num_shards(257) == 2
    assert orc.Oracle.num_shards(1 << 12) == 16
    assert orc.Oracle.num_shards(214_000_000) == 1 << 20


# --- (d) invariants ------------------------------------------------------------------------------

@pytest.mark.parametrize("alphabet,k,s,m,h", [("dna", 31, 28, 16, 4), ("dna", 21, 19, 11, 8),
                                              ("protein", 10, 8, 6, 4), ("triplet", 7, 5, 4, 12),
                                              ("custom_xyz", 12, 10, 6, 16)])
def test_naive_and_fast_forms_agree(alphabet, k, s, m, h):
    o = make_oracle(k, s, m, h, alphabet)
    seq = random_bytes_for(alphabet, 60_000, seed=k + h, invalid_rate=0.002)
    n = o.num_shards(60_000 * 4)
    w1 = o.new_filter(n)
    o.insert(seq, w1)
    w2 = o.new_filter(n)
    o.mt_insert(seq, w2, threads=3)
    assert (w1 == w2).all()
    q = random_bytes_for(alphabet, 30_000, seed=99, invalid_rate=0.002) + seq[:30_000]
    assert (o.contains(q, w1) == o.mt_contains(q, w1, threads=4)).all()


def test_insert_order_and_split_independence():
    o = make_oracle(31, 28, 16, 4)
    seq = random_dna(40_000, seed=5, n_rate=0.001)
    n = o.num_shards(40_000 * 8)
    whole = o.new_filter(n)
    o.insert(seq, whole)
    parts = o.new_filter(n)
    cut = 17_003
    o.insert(seq[cut:], parts)          # reverse order, overlapping by k-1
    o.insert(seq[:cut + 30], parts)
    assert (whole == parts).all()


def test_lowercase_matches_uppercase():
    o = make_oracle(31, 28, 16, 4)
    seq = random_dna(10_000, seed=11)
    n = o.num_shards(80_000)
    a = o.new_filter(n)
    o.insert(seq, a)
    b = o.new_filter(n)
    o.insert(seq.lower(), b)
    assert (a == b).all()


def test_dense_path_reads_invalid_symbols_as_their_masked_code():
    # dense_packed.hpp:183-185: 'N' is stored as 0xFF & 3 = 3 (the code of G),
    # so the dense path inserts the windows of the sequence with N read as G.
    o = make_oracle(31, 28, 16, 4)
    seq = random_dna(5_000, seed=2, n_rate=0.01)
    n = o.num_shards(40_000)
    wd = o.new_filter(n)
    o.insert_dense(o.pack_dense(seq), len(seq), wd)
    wg = o.new_filter(n)
    o.insert(seq.replace(b"N", b"G"), wg)
    assert (wd == wg).all()


def test_random_negatives_fpr_is_small():
    # 21.4 bits per inserted k-mer exactly (2^22 bits / 195,995 windows). On a
    # uniform random sequence shard loads are Poisson (about 1.4 super-k-mers
    # per shard), and a negative needs 16 set bits in its shard: summing
    # P(load r) * fill(r)^16 over r gives about 1e-5. (The reference's 0.85%
    # at this density is measured on a real genome, whose skewed minimizer
    # distribution overloads some shards.)
    o = make_oracle(31, 28, 16, 4)
    seq = random_dna(195_995 + 30, seed=21)
    w = o.new_filter(o.num_shards(1 << 22))
    o.mt_insert(seq, w)
    fpr = o.mt_contains(random_dna(1_000_000, seed=22), w).mean()
    assert 1e-6 < fpr < 1e-3, fpr


# --- filter images and load_factor (oracle/image_oracle.py) -------------------------------------

def test_splitmix64_finaliser_matches_published_outputs():
    from oracle import image_oracle as im

    # SplitMix64 from seed 0 (Vigna's splitmix64.c): the published first outputs
    assert im.splitmix64_sequence(0, 5) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                            0x06C45D188009454F, 0xF88BB8A8724C81EC,
                                            0x1B39896A51A8749B]


def test_image_checksum_definition():
    from oracle import image_oracle as im

    assert im.words_checksum(np.zeros(0, np.uint64)) == 0
    # word i = 0: mix(i * golden) is SplitMix64's output i from seed 0
    seq = im.splitmix64_sequence(0, 4)
    assert im.words_checksum(np.zeros(5, np.uint64)) == sum(seq) + int(im.splitmix64_mix(np.zeros(1, np.uint64))[0]) & MASK64
    # the index enters: swapping two different words changes the sum
    w = np.array([1, 2, 3], dtype=np.uint64)
    assert im.words_checksum(w) != im.words_checksum(w[::-1].copy())
    # chunking does not change the sum
    r = np.random.default_rng(1).integers(0, 1 << 63, 1000, dtype=np.uint64)
    assert im.words_checksum(r) == im.words_checksum(r, chunk=7)


def test_load_factor_restatement():
    from oracle import image_oracle as im

    assert im.load_factor(np.zeros(64, np.uint64), 4096) == 0.0
    assert im.load_factor(np.full(64, MASK64, np.uint64), 4096) == 1.0
    one = np.zeros(1 << 10, np.uint64)
    one[77] = 1 << 13
    assert im.load_factor(one, 1 << 16) == 2.0 ** -16
    r = np.random.default_rng(2).integers(0, 1 << 63, 4096, dtype=np.uint64)
    assert im.popcount_words(r) == sum(bin(int(x)).count("1") for x in r)
    # float32 rounding of the quotient, as the reference's float division
    assert im.load_factor(r, 4096 * 64) == float(np.float32(im.popcount_words(r)) / np.float32(4096 * 64))


def test_normalize_records_against_reference_batches():
    c = KNOWN["device_record_batch"]
    seq = (c["record1"] + c["separator"] + c["record2"]).encode()
    norm = orc.normalize_records(seq, c["ranges"], ord(c["separator"]))
    assert norm == seq
    o = oracle_for(c["config"])
    w = o.new_filter(o.num_shards(c["filter_bits"]))
    o.insert(c["record1"], w)
    o.insert(c["record2"], w)
    assert o.contains(norm, w).tolist() == c["batch_hits"]
    for case in KNOWN["normalized_record_batch"]:
        if "total_valid_kmers" not in case:
            continue
        o = oracle_for(case["config"])
        norm = orc.normalize_records(case["sequence"], case["ranges"], o.separator, o.symbol_width)
        assert len(norm) > case["min_normalized_size_exclusive"]
        assert o.count(norm, 1)[0] == case["total_valid_kmers"]
    # width 3: padding to the symbol width, then one separator symbol
    assert orc.normalize_records(b"ACGTA", [(0, 4), (4, 1)], ord("N"), 3) == b"ACGTNNNNNA"
    # empty leading records write no separator
    assert orc.normalize_records(b"ACGT", [(0, 0), (0, 4)], ord("N")) == b"ACGT"
