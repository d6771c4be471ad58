"""GPU parity of the runtime-span lane-strip kernels (strip_wide.hpp).

One kernel set per (alphabet, h) serves every DNA configuration with k <= 32,
k - s + 1 <= 16 and k - m + 1 <= 24 (the reference's DNA parameter sweep,
benchmarks/param_sweep.cu:263-530) and every protein configuration with k <= 12
(its protein grid, :534-591). Filter words and answers must equal the CPU
oracle's: sweep corners (findere span 1..16, minimizer span 1..24, every hash
count), small k, protein, the atomic and the sorted insert, dense packed input,
the batched record form (strided and work-queue kernels), tile seams, the
chunked query of minimizer spans up to 3 (more runs per tile than LDS slots),
and the compile-time lane strips' configurations run through the runtime-span
kernels.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, make_oracle, random_bytes_for
from tests.test_gpu_batched import (device, device_offsets, expected_hits, lognormal_lengths,
                                    mutate_every_other, offsets_of, oracle_words, symbols_for)
from tests.test_gpu_parity import dev

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


WIDE = [
    # DNA k = 31 sweep corners: (k, s, m, h), findere span f = k - s + 1, minimizer span w = k - m + 1
    ("dna", 31, 16, 8, 4),    # f 16, w 24 (both limits)
    ("dna", 31, 16, 31, 16),  # f 16, w 1
    ("dna", 31, 31, 8, 8),    # f 1, w 24
    ("dna", 31, 17, 10, 12),  # f 15, w 22
    ("dna", 31, 23, 14, 4),   # f 9, w 18
    ("dna", 31, 24, 16, 12),  # f 8, w 16
    ("dna", 31, 20, 12, 8),   # f 12, w 20
    ("dna", 31, 27, 30, 16),  # f 5, w 2
    ("dna", 31, 29, 22, 4),   # f 3, w 10
    ("dna", 31, 25, 18, 8),   # f 7, w 14
    ("dna", 31, 16, 16, 4),   # f 16, w 16
    ("dna", 31, 19, 26, 16),  # f 13, w 6
    ("dna", 31, 22, 29, 8),   # f 10, w 3 (chunked query: more runs than slots)
    # other k
    ("dna", 32, 20, 32, 8),
    ("dna", 32, 17, 9, 4),
    ("dna", 27, 25, 13, 4),
    ("dna", 31, 28, 18, 16),
    ("dna", 17, 15, 9, 4),
    ("dna", 8, 5, 3, 4),
    ("dna", 2, 1, 1, 4),
    ("dna", 1, 1, 1, 8),
    # protein (k <= 12)
    ("protein", 12, 6, 4, 16),   # f 7, w 9
    ("protein", 12, 12, 12, 4),  # f 1, w 1
    ("protein", 12, 9, 8, 8),
    ("protein", 12, 7, 10, 12),
    ("protein", 12, 8, 5, 16),
    ("protein", 11, 6, 4, 4),
    ("protein", 7, 3, 2, 4),
]


def _ids(c):
    return f"{c[0]}-k{c[1]}s{c[2]}m{c[3]}h{c[4]}"


def _check(f, alphabet, k, s, m, h, seq, query):
    torch = _torch()
    o = make_oracle(k, s, m, h, alphabet)
    w = o.new_filter(f.num_shards())
    o.mt_insert(seq, w)
    got = f.words()
    bad = np.nonzero(got != w)[0]
    assert bad.size == 0, f"{bad.size} filter words differ (first shard {bad[0] // 4})"
    out = torch.empty(max(1, o.kmer_count(len(query))), dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(dev(query), out)
    torch.cuda.synchronize()
    want = o.mt_contains(query, w)
    hits = out.cpu().numpy()[: want.size]
    bad = np.nonzero(hits != want)[0]
    assert bad.size == 0, f"{bad.size} answers differ (first window {bad[0]})"
    return want


def _inputs(alphabet, n, seed):
    seq = random_bytes_for(alphabet, n, seed, 0.003)
    query = random_bytes_for(alphabet, n // 2, seed + 7, 0.003) + seq[: n // 2]
    return seq, query


@pytest.mark.parametrize("path", ["atomic", "sorted"])
@pytest.mark.parametrize("case", WIDE, ids=_ids)
def test_wide_parity(gpu, case, path):
    import cusbf_amd as cs

    torch = _torch()
    alphabet, k, s, m, h = case
    seq, query = _inputs(alphabet, 150_001, seed=k * 131 + s * 7 + m + h)
    f = cs.Filter(make_config(k, s, m, h, alphabet), 8 * len(seq))
    assert f.kernel_family() == "lane_strip_wide"
    f.set_insert_path(cs.INSERT_ATOMIC if path == "atomic" else cs.INSERT_SORTED)
    f.insert_sequence_async(dev(seq))
    torch.cuda.synchronize()
    assert f.insert_path().last_path == (cs.INSERT_ATOMIC if path == "atomic" else cs.INSERT_SORTED)
    want = _check(f, alphabet, k, s, m, h, seq, query)
    assert 0 < want.sum() < want.size  # present and absent windows both tested


@pytest.mark.parametrize("case", [("dna", 31, 28, 16, 4), ("dna", 31, 30, 16, 4), ("dna", 31, 31, 31, 4),
                                  ("dna", 31, 28, 24, 4), ("dna", 31, 28, 16, 16), ("dna", 5, 4, 3, 4),
                                  ("dna", 3, 2, 2, 4), ("protein", 10, 8, 6, 4), ("protein", 12, 10, 6, 4),
                                  ("protein", 5, 4, 3, 4)], ids=_ids)
def test_wide_equals_compiled_lane_strips(gpu, monkeypatch, case):
    # the configurations with compile-time kernels, forced through the runtime spans
    import cusbf_amd as cs

    torch = _torch()
    alphabet, k, s, m, h = case
    seq, query = _inputs(alphabet, 200_003, seed=k + s + m + h)
    fixed = cs.Filter(make_config(k, s, m, h, alphabet), 8 * len(seq))
    monkeypatch.setenv("CUSBF_FORCE_WIDE", "1")
    wide = cs.Filter(make_config(k, s, m, h, alphabet), 8 * len(seq))
    assert fixed.kernel_family() == "lane_strip" and wide.kernel_family() == "lane_strip_wide"
    fixed.insert_sequence_async(dev(seq))
    wide.insert_sequence_async(dev(seq))
    torch.cuda.synchronize()
    assert (fixed.words() == wide.words()).all()
    _check(wide, alphabet, k, s, m, h, seq, query)


@pytest.mark.parametrize("n", [31, 32, 46, 47, 991 + 30, 992 + 30, 993 + 30, 3968 + 30, 65_536 + 29])
def test_wide_tile_seams(gpu, n):
    import cusbf_amd as cs

    torch = _torch()
    seq, query = _inputs("dna", n, seed=n)
    f = cs.Filter(make_config(31, 20, 10, 8), 64 * max(n, 1))
    assert f.kernel_family() == "lane_strip_wide"
    f.insert_sequence_async(dev(seq))
    torch.cuda.synchronize()
    _check(f, "dna", 31, 20, 10, 8, seq, seq + query)


@pytest.mark.parametrize("case", [("dna", 31, 16, 8, 4), ("dna", 31, 22, 20, 12), ("dna", 31, 24, 31, 8)],
                         ids=_ids)
def test_wide_dense_parity(gpu, case):
    import cusbf_amd as cs

    torch = _torch()
    alphabet, k, s, m, h = case
    cfg = make_config(k, s, m, h, alphabet)
    o = make_oracle(k, s, m, h, alphabet)
    seq = random_bytes_for(alphabet, 120_003, seed=k + h, invalid_rate=0.002)
    packed = o.pack_dense(seq)
    d_words = torch.from_numpy(packed.view(np.int64)).cuda()
    f = cs.Filter(cfg, 8 * len(seq))
    assert f.kernel_family() == "lane_strip_wide"
    f.insert_dense_packed_async(d_words, len(seq))
    torch.cuda.synchronize()
    w = o.new_filter(f.num_shards())
    o.insert_dense(packed, len(seq), w)
    assert (f.words() == w).all()
    out = torch.empty(len(seq) - k + 1, dtype=torch.uint8, device="cuda")
    f.contains_dense_packed_async(d_words, len(seq), out)
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == o.contains_dense(packed, len(seq), w)).all()


@pytest.mark.parametrize("claim", [0, 4])
@pytest.mark.parametrize("sorted_path", [False, True])
@pytest.mark.parametrize("case", [("dna", 31, 18, 12, 8), ("dna", 17, 15, 9, 4), ("protein", 12, 7, 5, 4)],
                         ids=_ids)
def test_wide_batched(gpu, case, claim, sorted_path):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    alphabet, k, s, m, h = case
    lens = lognormal_lengths(1500, 300, 1.5, seed=k * 100 + s)
    lens[:5] = [0, 1, k - 1, k, k + 1]  # empty, shorter than k, exactly k
    off = offsets_of(lens)
    data = symbols_for(alphabet, int(off[-1]), seed=7 + k)
    query = mutate_every_other(alphabet, data, off, seed=99 + k)
    f = cs.Filter(make_config(k, s, m, h, alphabet), max(1 << 16, int(off[-1]) * 16))
    assert f.kernel_family() == "lane_strip_wide"
    f.set_insert_path(cs.INSERT_SORTED if sorted_path else cs.INSERT_ATOMIC)
    for op in (bt.INSERT, bt.INSERT_SORTED, bt.CONTAINS):
        for bucket in range(bt.BUCKETS):
            bt.set_tuning(f, bucket, op, bt.BatchTuning(3, 4, claim))
    bt.insert_batched_async(f, device(data), device_offsets(off))
    bt.batched_status(f)
    o = make_oracle(k, s, m, h, alphabet)
    w = oracle_words(o, data, off, alphabet, f.num_shards())
    assert (f.words() == w).all()
    n_win = int(off[-1]) - k + 1
    out = torch.full((n_win + 7,), 0xEE, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(f, device(query), device_offsets(off), out)
    bt.batched_status(f)
    got = out.cpu().numpy()
    exp = expected_hits(o, w, query, off, k, alphabet)
    assert np.array_equal(got[:n_win], exp)
    assert (got[n_win:] == 0xEE).all()
    assert 0 < exp.sum() < n_win


# The reference's parameter grids (benchmarks/param_sweep.cu:263-591): DNA k = 31 with
# s = 16..31 and m = 8, 10, .., 30, 31, protein k = 12 with s = 6..12 and m = 4, 6, .., 12;
# each (s, m) point once, the hash count cycling through 4, 8, 12 and 16.
DNA_GRID = [(31, s, m, (4, 8, 12, 16)[(s + m) % 4]) for s in range(16, 32)
            for m in list(range(8, 31, 2)) + [31]]
PROTEIN_GRID = [(12, s, m, (4, 8, 12, 16)[(s + m) % 4]) for s in range(6, 13) for m in range(4, 13, 2)]


@pytest.mark.parametrize("alphabet,grid", [("dna", DNA_GRID), ("protein", PROTEIN_GRID)])
def test_reference_sweep_grids_equal_oracle(gpu, alphabet, grid):
    """Every (s, m) point of the reference's sweep grids on the dispatched kernels (runtime-span
    or compiled lane strips): words of a seeded insert and answers of a query half present,
    half fresh, equal to the oracle's."""
    import cusbf_amd as cs

    torch = _torch()
    bad = []
    for i, (k, s, m, h) in enumerate(grid):
        seq, query = _inputs(alphabet, 12_001, seed=1000 + i)
        f = cs.Filter(make_config(k, s, m, h, alphabet), 8 * len(seq))
        if f.kernel_family() == "generic":
            bad.append((k, s, m, h, "generic kernels"))
            continue
        f.insert_sequence_async(dev(seq))
        torch.cuda.synchronize()
        o = make_oracle(k, s, m, h, alphabet)
        w = o.new_filter(f.num_shards())
        o.insert(seq, w)
        out = torch.empty(o.kmer_count(len(query)), dtype=torch.uint8, device="cuda")
        f.contains_sequence_async(dev(query), out)
        torch.cuda.synchronize()
        if not (f.words() == w).all():
            bad.append((k, s, m, h, "words"))
        elif not (out.cpu().numpy() == o.contains(query, w)).all():
            bad.append((k, s, m, h, "answers"))
        f.close()
    assert not bad, bad
