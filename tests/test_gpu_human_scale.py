"""The human-scale workload at full size, against the oracle.

Seeded 3.1 Gbp genome (24 chromosomes joined by separators, 45% repeat copies)
inserted by the sorted insert into a 2^36-bit filter; then

  * all 2^30 filter words equal the multithreaded oracle's, and load_factor
    equals the oracle's popcount / filter bits;
  * the headline query -- 100 M seeded 150 bp reads joined by separators, half
    from the genome at 1% substitutions, half foreign -- gives exactly the
    oracle's ~15.1 G answer bytes;
  * serving: a capped request set in the serving configuration's mix (sizes
    log-uniform over 1..1M reads; 150 bp reads at 1% substitutions or 10-50 kbp
    long reads at 5% mixed substitutions and indels) gives, for every request,
    the oracle's answers and report on its normalised batch, alone and when the
    requests run together through the continuous-batching server.

Tolerance: none (integer answers, bit-exact).
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import make_config, make_oracle

pytestmark = pytest.mark.gpu

K, S, M, H = 31, 28, 16, 4


@pytest.fixture(scope="module")
def human(gpu):
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    torch.cuda.empty_cache()  # earlier tests' cached blocks would shrink the free memory
    wl = W.WORKLOADS["human"]
    genome, _ = W.generate_genome(wl.genome)
    n = genome.numel()
    f = cs.Filter(make_config(K, S, M, H), int(wl.bits_per_kmer * (n - K + 1)))
    assert f.num_shards() == 1 << 28
    f.set_insert_path(cs.INSERT_SORTED)
    f.insert_sequence_async(genome)
    torch.cuda.synchronize()
    assert f.insert_path().last_path == cs.INSERT_SORTED
    o = make_oracle(K, S, M, H)
    w = o.new_filter(f.num_shards())
    o.mt_insert(genome.cpu().numpy(), w)
    yield {"filter": f, "genome": genome, "words": w, "oracle": o, "wl": wl}
    f.close()
    del genome
    torch.cuda.empty_cache()


def test_words_and_load_factor_equal_oracle(human):
    """Every word of the 8 GiB filter, and load_factor, against the oracle."""
    import torch

    import cusbf_amd as cs
    from oracle import image_oracle as im

    f, genome, w = human["filter"], human["genome"], human["words"]
    assert np.array_equal(f.words(), w)
    assert f.load_factor() == im.load_factor(w, f.filter_bits())
    # self-query: every valid window is present
    out = torch.empty(genome.numel() - K + 1, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(genome, out)
    valid, _ = f.count_windows_runs(genome)
    assert cs.count_positive(out) == valid


def test_headline_query_answers_equal_oracle(human):
    """The 100 M-read query of the headline workload: all answer bytes equal the
    oracle's, positives and negatives both present."""
    import torch

    from cusbf_amd import workloads as W

    f, genome, w, o, wl = human["filter"], human["genome"], human["words"], human["oracle"], human["wl"]
    reads = W.generate_reads(wl.reads, genome)
    assert reads.numel() == wl.reads.n_reads * 151 - 1
    out = torch.empty(reads.numel() - K + 1, dtype=torch.uint8, device="cuda")
    f.contains_sequence_async(reads, out)
    got = out.cpu().numpy()
    host_reads = reads.cpu().numpy()
    del out, reads
    torch.cuda.empty_cache()
    want = o.mt_contains(host_reads, w)
    del host_reads
    assert got.size == want.size
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} answers differ (first window {bad[0]})"
    positives = int(np.count_nonzero(want))
    assert 0 < positives < want.size


def _oracle_request(o, w, seq, ranges):
    """(answers of the normalised batch, records, bases, valid windows, positives)."""
    norm = orc.normalize_records(seq, ranges, ord("N"))
    want = o.mt_contains(norm, w)
    valid, _ = o.count(norm, w.size // 4)
    return want, (len(ranges), int(ranges[:, 1].sum()), valid, int(np.count_nonzero(want)))


def test_serving_mix_equals_oracle_alone_and_together(human):
    """60 requests in the serving mix on the resident 2^36-bit filter. Each request's
    answers and FastxQueryReport equal the oracle's on its normalised batch, when
    queried alone and when all run together through the server (4 slots, 16 GB
    staging, submitted from three threads)."""
    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    f, genome, w, o, wl = human["filter"], human["genome"], human["words"], human["oracle"], human["wl"]
    pool = W.ReadSpec(400_000, 150, 0.5, 0.01, seed=61)
    short = W.generate_reads(pool, genome).cpu().numpy()
    d_long, long_off = W.generate_long_reads(genome, 2_000, error=0.05, seed=62)
    long_pool = d_long.cpu().numpy()
    del d_long
    reqs = W.serving_requests(short, 150, long_pool, long_off, 60, seed=63)
    assert any(r[1][0, 1] == 150 for r in reqs) and any(r[1][0, 1] >= 10_000 for r in reqs)
    expected = [_oracle_request(o, w, seq, rg) for seq, rg in reqs]
    for i, (seq, rg) in enumerate(reqs):
        hits, rep = cs.query_record_batch_hits(f, seq, rg)
        want, counts = expected[i]
        assert np.array_equal(hits, want), f"request {i} alone"
        assert (rep.records_queried, rep.queried_bases, rep.queried_kmers,
                rep.positive_kmers) == counts, f"request {i} alone"
    with cs.Server(f, staging_bytes=16 << 30, slots=4) as srv:
        pend = [None] * len(reqs)

        def submit(idxs):
            for i in idxs:
                pend[i] = srv.submit(*reqs[i])

        threads = [threading.Thread(target=submit, args=(range(t, len(reqs), 3),)) for t in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for i, p in enumerate(pend):
            hits, rep, _ = srv.wait(p)
            want, counts = expected[i]
            assert np.array_equal(hits, want), f"request {i} in the mix"
            assert (rep.records_queried, rep.queried_bases, rep.queried_kmers,
                    rep.positive_kmers) == counts, f"request {i} in the mix"
        st = srv.stats()
    assert st["requests"] == len(reqs)
