"""Serving form: continuous batching returns, for every request, exactly the answer
the request gets alone (query_record_batch with hits, filter.cuh:566-582,
1007-1131). Tolerance: none (bit-exact hits and equal reports)."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from cusbf_amd.serve import normalized_size, ranges_array
from tests.helpers import make_config, random_dna

pytestmark = pytest.mark.gpu


def _requests(genome: bytes, n: int, seed: int):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        if rng.random() < 0.5:  # short reads, some from the genome
            count = int(np.exp(rng.uniform(0, np.log(3000))))
            parts = []
            for _ in range(count):
                if rng.random() < 0.5:
                    st = int(rng.integers(0, len(genome) - 150))
                    parts.append(genome[st:st + 150])
                else:
                    parts.append(random_dna(150, int(rng.integers(1 << 30))))
            seq = b"".join(parts)
            ranges = [(150 * j, 150) for j in range(count)]
        else:  # a few long records, with N runs
            lens = rng.integers(31, 20000, int(rng.integers(1, 6)))
            st = int(rng.integers(0, len(genome) - int(lens.sum())))
            seq = bytearray(genome[st:st + int(lens.sum())])
            for p in rng.integers(0, len(seq), 3):
                seq[p] = ord("N")
            seq = bytes(seq)
            offs = np.concatenate([[0], np.cumsum(lens)[:-1]])
            ranges = [(int(o), int(l)) for o, l in zip(offs, lens)]
        reqs.append((seq, ranges))
    return reqs


@pytest.fixture(scope="module")
def loaded(gpu):
    import cusbf_amd as cs

    genome = random_dna(3_000_000, seed=123, n_rate=0.0005)
    f = cs.Filter(make_config(31, 28, 16, 4), int(21.4 * len(genome)))
    f.insert_sequence(genome)
    return f, genome


def test_each_request_same_alone_and_in_the_mix(loaded):
    import cusbf_amd as cs

    f, genome = loaded
    reqs = _requests(genome, 60, seed=7)
    alone = [cs.query_record_batch_hits(f, s, r) for s, r in reqs]
    # small slots so requests queue, coalesce and wait for slots to free up
    with cs.Server(f, staging_bytes=8 << 20, slots=3) as srv:
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
            hits, rep, lat = srv.wait(p)
            want_hits, want_rep = alone[i]
            assert hits.shape == want_hits.shape
            assert (hits == want_hits).all(), f"request {i} differs in the mix"
            assert rep == want_rep
            assert lat > 0
        st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["batches"] < len(reqs)  # requests were coalesced
    assert st["max_requests_per_batch"] > 1


def _gapped_requests(genome: bytes, n: int, seed: int):
    """Requests whose records sit one byte apart in their buffer (reads or long records
    with a gap byte between them, which the server copies as one span and patches with
    separators); the gap byte is sometimes a valid symbol, so a missing separator would
    change the answers of the windows across it."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        count = int(np.exp(rng.uniform(0, np.log(5000))))
        lens = (np.full(count, 150) if rng.random() < 0.5
                else rng.integers(31, 30000, max(1, count // 200)))
        parts, ranges, off = [], [], 0
        for ln in lens.tolist():
            st = int(rng.integers(0, len(genome) - ln))
            rec = genome[st:st + ln] if rng.random() < 0.7 else random_dna(ln, int(rng.integers(1 << 30)))
            if parts:
                parts.append(b"ACGTN"[int(rng.integers(0, 5))].to_bytes(1, "little"))
                off += 1
            parts.append(rec)
            ranges.append((off, ln))
            off += ln
        reqs.append((b"".join(parts), ranges))
    return reqs


def test_span_layout_requests_alone_and_in_the_mix(loaded):
    import cusbf_amd as cs

    f, genome = loaded
    reqs = _gapped_requests(genome, 40, seed=11) + _requests(genome, 10, seed=12)
    alone = [cs.query_record_batch_hits(f, s, r) for s, r in reqs]
    with cs.Server(f, staging_bytes=64 << 20, slots=3) as srv:
        pend = [srv.submit(*r) for r in reqs]
        for i, p in enumerate(pend):
            hits, rep, _ = srv.wait(p)
            want_hits, want_rep = alone[i]
            assert np.array_equal(hits, want_hits), f"request {i} differs in the mix"
            assert rep == want_rep


def test_hits_match_plain_query_of_normalized_request(loaded):
    import cusbf_amd as cs

    f, genome = loaded
    seq, ranges = _requests(genome, 1, seed=3)[0]
    hits, rep = cs.query_record_batch_hits(f, seq, ranges)
    norm, meta = cs.normalize_record_batch(f.config, seq, ranges)
    assert (hits == f.contains_sequence(norm)).all()
    assert rep.queried_kmers == sum(m.valid_kmers for m in meta)
    arr, n = ranges_array(ranges)
    assert normalized_size(arr, n, 1) == len(norm)


def test_oversized_request_is_a_resource_error(loaded):
    import cusbf_amd as cs

    f, genome = loaded
    with cs.Server(f, staging_bytes=1 << 20, slots=2) as srv:
        big = genome[:600_000]
        with pytest.raises(cs.CusbfError) as e:
            srv.submit(big, [(0, len(big))])
        assert e.value.category == cs.ErrorCategory.resource
        p = srv.submit(big[:1000], [(0, 1000)])
        hits, rep, _ = srv.wait(p)
        assert hits.all() or rep.positive_kmers <= rep.queried_kmers
