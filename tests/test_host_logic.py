"""Host-side logic and the C ABI surface (CPU only, no kernel launches).

- the library loads and exports every entry point include/cusbf_hip.h declares;
- configuration validation mirrors config.cuh:74-117 with its messages;
- record-batch normalisation and the query layout reproduce the reference's
  tests/test_normalized_record_batch.cpp and tests/test_query_layout.cpp;
- host dense packing equals the oracle's.
"""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import cusbf_amd as cs
from cusbf_amd import _lib
from cusbf_amd.normalize import NormalizedRecord, query_layout
from tests.helpers import KNOWN, config_for, make_config, make_oracle, random_bytes_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cusbf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cusbf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/cusbf_hip.h but not exported"
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound
    assert lib.cusbf_abi_version() == 1


def test_config_validation_messages():
    with pytest.raises(cs.CusbfError) as e:
        cs.Config(31, 28, 16, 6)
    assert e.value.category == cs.ErrorCategory.invalid_argument
    assert e.value.message == "Hash count must distribute evenly across shard words"
    with pytest.raises(cs.CusbfError) as e:
        cs.Config(33, 28, 16, 4)
    assert e.value.message == "k-mer must fit in one packed uint64_t"
    with pytest.raises(cs.CusbfError) as e:
        cs.Config(31, 32, 16, 4)
    assert e.value.message == "s must satisfy 0 < s <= k"
    with pytest.raises(cs.CusbfError):
        cs.Config(13, 8, 6, 4, alphabet=cs.ProteinAlphabet)  # 13 * 5 bits > 64
    c = cs.Config(12, 8, 5, 4, alphabet=cs.ProteinAlphabet)  # test_support.cuh:52
    assert (c.symbol_bits, c.symbol_width) == (5, 1)
    assert cs.Config(3, 2, 2, 4, alphabet=cs.DnaTripletAlphabet).symbol_bits == 6
    assert cs.Config(5, 4, 3, 4).symbol_bits == 2


def test_custom_alphabet_separator_must_be_invalid():
    with pytest.raises(ValueError):
        cs.Alphabet.custom("bad", 3, 7, "x", b"xyz", lambda b: {120: 0, 121: 1, 122: 2}.get(b, 7))


@pytest.mark.parametrize("case", KNOWN["query_layout"], ids=lambda c: c["source"])
def test_query_layout_known_answers(case):
    cfg = config_for(case["config"])
    recs = [NormalizedRecord(*r) for r in case["records"]]
    layout, total = query_layout(cfg, recs)
    assert total == case["total_hit_count"]
    assert [r.hit_offset for r in layout] == case["hit_offsets"]
    assert [r.hit_count for r in layout] == case["hit_counts"]
    assert [r.valid_kmers for r in layout] == [r[3] for r in case["records"]]


def test_normalized_record_batch_known_answers():
    ok, bad = KNOWN["normalized_record_batch"]
    cfg = config_for(ok["config"])
    norm, recs = cs.normalize_record_batch(cfg, ok["sequence"], ok["ranges"])
    assert len(norm) > ok["min_normalized_size_exclusive"]
    assert sum(r.valid_kmers for r in recs) == ok["total_valid_kmers"]
    assert [r.record_index for r in recs] == ok["record_indices"]
    assert [r.input_offset for r in recs] == ok["input_offsets"]
    with pytest.raises(cs.CusbfError) as e:
        cs.normalize_record_batch(config_for(bad["config"]), bad["sequence"], bad["ranges"])
    assert int(e.value.category) == bad["error_category"]


def test_normalization_separators_and_triplet_padding():
    cfg = cs.Config(3, 2, 2, 4, alphabet=cs.DnaTripletAlphabet)
    norm, recs = cs.normalize_record_batch(cfg, "ACGTTTGGGCCC", [(0, 6), (6, 6)])
    assert norm == b"ACGTTTNNNGGGCCC"
    with pytest.raises(cs.CusbfError) as e:
        cs.normalize_record_batch(cfg, "ACGTTTGGGCCC", [(0, 4)])
    assert "symbol width" in e.value.message
    with pytest.raises(cs.CusbfError) as e:
        cs.normalize_record_batch(cfg, "ACGTTT", [(0, 9)])
    assert e.value.message == "record batch range exceeds the source sequence"


def test_normalized_valid_kmers_match_oracle_windows():
    cfg = make_config(31, 28, 16, 4)
    o = make_oracle(31, 28, 16, 4)
    seq = random_bytes_for("dna", 5000, seed=3, invalid_rate=0.01)
    norm, recs = cs.normalize_record_batch(cfg, seq, [(0, 2000), (2000, 3000)])
    valid, _ = o.count(seq[:2000], 1)
    assert recs[0].valid_kmers == valid
    valid, _ = o.count(seq[2000:], 1)
    assert recs[1].valid_kmers == valid


@pytest.mark.parametrize("alphabet,k", [("dna", 31), ("protein", 10), ("triplet", 7),
                                        ("custom_xyz", 12)])
def test_host_dense_packing_matches_oracle(alphabet, k):
    cfg = make_config(k, k - 2, k // 2, 4, alphabet)
    o = make_oracle(k, k - 2, k // 2, 4, alphabet)
    seq = random_bytes_for(alphabet, 10_007, seed=k, invalid_rate=0.01)
    assert (cs.pack_dense_sequence(cfg, seq) == o.pack_dense(seq)).all()


def test_result_wrapper():
    r = cs.Result.of(cs.Config, 31, 28, 16, 5)
    assert not r and r.error().category == cs.ErrorCategory.invalid_argument
    r = cs.Result.of(cs.Config, 31, 28, 16, 4)
    assert r and r.value().k == 31
