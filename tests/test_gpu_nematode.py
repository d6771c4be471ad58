"""The nematode-scale workload at full size, against the oracle.

Seeded 100 Mbp genome (35% GC, 6 chromosomes of 12-21 Mbp joined by
separators, 5% tandem and interspersed repeats at 5% divergence), k=31, s=28,
m=16, h=4 at 85.7 bits per k-mer (2^33 bits after rounding), inserted by the
library's own path choice; 20 M seeded 150 bp reads (half from the genome at
0.5% substitutions, half uniform random) queried. All 2^27 filter words and all
~2.4 G answer bytes equal the multithreaded oracle's.

Beside it, the reference's FPR densities (README.md:59-63; protocol
benchmarks/fpr_fastx_sweep.cu: the genome inserted, uniform random k-mers
queried, m = 21): 2^31, 2^33 and 2^35 bits (21.4 / 85.7 / 342.6 bits per k-mer
for this genome), where the answers on a seeded uniform query equal the
oracle's, so the empirical FPR is the reference algorithm's.

Tolerance: none (integer answers, bit-exact).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, make_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nematode(gpu):
    import torch

    from cusbf_amd import workloads as W

    torch.cuda.empty_cache()
    wl = W.WORKLOADS["nematode"]
    genome, lens = W.generate_genome(wl.genome)
    assert genome.numel() == 100_000_000 + 5 and lens.size == 6
    assert lens.min() >= 12_000_000 and lens.max() <= 21_000_000
    yield {"genome": genome, "host": genome.cpu().numpy(), "wl": wl}
    del genome
    torch.cuda.empty_cache()


def test_nematode_words_and_reads_equal_oracle(nematode):
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    genome, host, wl = nematode["genome"], nematode["host"], nematode["wl"]
    k, s, m, h = wl.k, wl.s, wl.m, wl.h
    f = cs.Filter(make_config(k, s, m, h), int(wl.bits_per_kmer * (genome.numel() - k + 1)))
    assert f.filter_bits() == 1 << 33
    f.insert_sequence_async(genome)
    torch.cuda.synchronize()
    o = make_oracle(k, s, m, h)
    w = o.new_filter(f.num_shards())
    o.mt_insert(host, w)
    assert np.array_equal(f.words(), w)
    reads = W.generate_reads(wl.reads, genome)
    assert reads.numel() == 20_000_000 * 151 - 1
    out = torch.empty(reads.numel() - k + 1, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(reads, out)
    got = out.cpu().numpy()
    want = o.mt_contains(reads.cpu().numpy(), w)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} answers differ (first window {bad[0]})"
    # genome-derived reads (first half) are mostly present, foreign reads mostly absent
    half = 10_000_000 * 151
    assert want[:half].mean() > 0.5 * 120 / 151 and want[half:].mean() < 0.01
    f.close()


@pytest.mark.parametrize("s,m,log2_bits", [(28, 21, 31), (28, 21, 33), (28, 21, 35), (31, 21, 31),
                                           (30, 21, 35), (28, 16, 35)])
def test_reference_fpr_densities_equal_oracle(nematode, s, m, log2_bits):
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    genome, host = nematode["genome"], nematode["host"]
    f = cs.Filter(make_config(31, s, m, 4), 1 << log2_bits)
    f.insert_sequence_async(genome)
    q = W.uniform_dna(20_000_030, seed=0xDEADBEEF + log2_bits)
    out = torch.empty(q.numel() - 30, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(q, out)
    got = out.cpu().numpy()
    o = make_oracle(31, s, m, 4)
    w = o.new_filter(f.num_shards())
    o.mt_insert(host, w)
    want = o.mt_contains(q.cpu().numpy(), w)
    assert np.array_equal(got, want)
    fpr = np.count_nonzero(want) / want.size
    assert 0 < fpr < 0.03, fpr
    f.close()
