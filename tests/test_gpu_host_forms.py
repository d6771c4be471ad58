"""GPU parity of the host-memory forms at sizes that stage through the pinned ring.

The host forms (filter.cuh:154-167 insert_sequence, :486-533 contains_sequence
into device or host memory, :226-240 / :267-312 the dense packed forms) copy
the caller's bytes through a ring of 32 MiB pinned buffers: host threads fill
one buffer while the DMA engine reads another, and answers come back the same
way. These inputs take several buffers and a ragged last one; the words and
answers must equal the device forms' on the same bytes and the oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, make_oracle, random_dna

pytestmark = pytest.mark.gpu

K, S, M, H = 31, 28, 16, 4


def _torch():
    import torch

    return torch


def _dev(b: bytes):
    torch = _torch()
    return torch.from_numpy(np.frombuffer(b, dtype=np.uint8).copy()).cuda()


def test_host_byte_forms_through_the_ring(gpu):
    import cusbf_amd as cs

    torch = _torch()
    cfg = make_config(K, S, M, H)
    seq = random_dna(150_000_003, 21, n_rate=0.001, lower_rate=0.01)  # 5 ring buffers
    bits = 1 << 30  # the words copy back through 4 buffers
    f = cs.Filter(cfg, bits)
    assert f.insert_sequence(seq) == len(seq) - K + 1
    g = cs.Filter(cfg, bits)
    g.insert_sequence_async(_dev(seq))
    torch.cuda.synchronize()
    o = make_oracle(K, S, M, H, "dna")
    w = o.new_filter(f.num_shards())
    o.mt_insert(seq, w)
    assert (g.words() == w).all()
    assert (f.words() == w).all()

    query = random_dna(70_000_001, 22, n_rate=0.001) + seq[:40_000_000]
    want = o.mt_contains(query, w)
    host = f.contains_sequence(query)
    bad = np.nonzero(host != want)[0]
    assert bad.size == 0, f"{bad.size} host answers differ (first window {bad[0]})"
    out = torch.full((want.size + 5,), 7, dtype=torch.uint8, device="cuda")
    assert f.contains_sequence(query, d_output=out) is None
    got = out.cpu().numpy()
    assert (got[: want.size] == want).all()
    assert (got[want.size:] == 7).all()
    # a second call reuses the ring buffers (the first call's copies have finished)
    assert (f.contains_sequence(query[:9_000_000]) == want[: 9_000_000 - K + 1]).all()


def test_host_dense_forms_through_the_ring(gpu):
    import cusbf_amd as cs

    torch = _torch()
    cfg = make_config(K, S, M, H)
    n = 300_000_017  # 9.4 M packed words: 3 ring buffers in; 300 M answers: 9 out
    seq = random_dna(n, 23)
    d_seq = _dev(seq)
    d_words = torch.empty((n + 31) // 32, dtype=torch.int64, device="cuda")
    cs.pack_dense_sequence_device(cfg, d_seq, n, d_words)
    torch.cuda.synchronize()
    words = d_words.cpu().numpy().view(np.uint64).copy()
    bits = 1 << 31
    f = cs.Filter(cfg, bits)
    assert f.insert_dense_packed(words, n) == n - K + 1
    g = cs.Filter(cfg, bits)
    g.insert_dense_packed_async(d_words, n)
    torch.cuda.synchronize()
    o = make_oracle(K, S, M, H, "dna")
    w = o.new_filter(f.num_shards())
    o.insert_dense(words, n, w)
    assert (g.words() == w).all()
    assert (f.words() == w).all()
    host = f.contains_dense_packed(words, n)
    dev_out = torch.empty(n - K + 1, dtype=torch.uint8, device="cuda")
    g.contains_dense_packed_async(d_words, n, dev_out)
    ref = dev_out.cpu().numpy()
    assert ref.all()  # every window of the inserted sequence is present
    assert (host == ref).all()
    out = torch.zeros(n - K + 1, dtype=torch.uint8, device="cuda")
    assert f.contains_dense_packed(words, n, d_output=out) is None
    assert torch.equal(out, dev_out)


@pytest.mark.parametrize("alphabet,cfg_t", [("dna", (31, 28, 16, 4)), ("triplet", (3, 2, 2, 4))])
def test_host_record_batch_through_the_ring(gpu, alphabet, cfg_t):
    """Host record batches (filter.cuh:338-364, 549-582, 930-990) of ~70 MB: the
    normalised buffer (records joined by separators padded to the symbol width) is
    produced chunk by chunk into the pinned ring by host threads, and the reports'
    valid windows are counted per record on the device. Words, reports and the
    per-window hits of the normalised batch equal the oracle's."""
    import cusbf_amd as cs
    from cusbf_amd.serve import query_record_batch_hits

    k, s, m, h = cfg_t
    width = 3 if alphabet == "triplet" else 1
    rng = np.random.default_rng(5)
    lens = np.clip(np.round(rng.lognormal(np.log(1500), 1.3, 20_000)), 0, 2_000_000).astype(np.int64)
    lens[:4] = [0, 2, 31, 1]
    gaps = rng.integers(0, 5, lens.size)
    if alphabet == "triplet":  # ranges align to whole codons (filter.cuh:429-457)
        lens, gaps = lens - lens % 3, 3 * gaps
    total = int(lens.sum() + gaps.sum())
    source = random_dna(total, 31, n_rate=0.002)
    ranges, at = [], 0
    for ln, gp in zip(lens.tolist(), gaps.tolist()):
        at += gp
        ranges.append((at, ln))
        at += ln
    records = [source[a:a + n] for a, n in ranges]
    cfg = make_config(k, s, m, h, alphabet)
    bits = 1 << 30
    f = cs.Filter(cfg, bits)
    rep = f.insert_record_batch(source, ranges)
    o = make_oracle(k, s, m, h, alphabet)
    w = o.new_filter(f.num_shards())
    # normalized_record_batch.hpp:208-232: a boundary (pad to the symbol width, then one
    # separator symbol) before every record that has bytes before it
    buf = bytearray()
    for r in records:
        if buf:
            buf += b"N" * ((-len(buf)) % width + width)
        buf += r
    joined = bytes(buf)
    o.mt_insert(joined, w)
    assert (f.words() == w).all()
    valid = sum(o.count(r, f.num_shards())[0] for r in records if len(r) >= k * width)
    assert (rep.records_indexed, rep.indexed_bases, rep.inserted_kmers) == (len(records), int(lens.sum()), valid)
    want = o.mt_contains(joined, w)
    q = f.query_record_batch(source, ranges)
    assert (q.records_queried, q.queried_bases, q.queried_kmers) == (len(records), int(lens.sum()), valid)
    assert q.positive_kmers == int(want.sum())
    hits, qh = query_record_batch_hits(f, source, ranges)
    assert (hits[: want.size] == want).all()
    assert (qh.queried_kmers, qh.positive_kmers) == (valid, int(want.sum()))


@pytest.mark.parametrize("n", [(4 << 20) - 1, 4 << 20, 32 << 20, (96 << 20) + 1])
def test_host_forms_at_ring_sizes(gpu, n):
    """Sizes at the ring's edges: below the 4 MiB direct-copy limit, exactly 4 MiB, one
    whole 32 MiB buffer, and one byte past three buffers (the ring wraps)."""
    import cusbf_amd as cs

    torch = _torch()
    cfg = make_config(K, S, M, H)
    seq = random_dna(n, n % 1000, n_rate=0.001)
    f = cs.Filter(cfg, 1 << 28)
    g = cs.Filter(cfg, 1 << 28)
    assert f.insert_sequence(seq) == n - K + 1
    g.insert_sequence_async(_dev(seq))
    torch.cuda.synchronize()
    assert (f.words() == g.words()).all()
    query = random_dna(n, n % 1000 + 1, n_rate=0.001)
    out = torch.empty(n - K + 1, dtype=torch.uint8, device="cuda")
    g.contains_sequence_async(_dev(query), out)
    want = out.cpu().numpy()
    assert (f.contains_sequence(query) == want).all()
    out2 = torch.full((n - K + 1,), 9, dtype=torch.uint8, device="cuda")
    f.contains_sequence(query, d_output=out2)
    assert torch.equal(out2, out)
