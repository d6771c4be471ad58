"""GPU parity of the sorted insert path (records, radix partition, LDS apply).

The sorted path must leave filter words bit-identical to the CPU oracle's and
to the atomic path's on the same input: plain sequences (DNA and protein front
ends, several k/s/m/h, tiles with more runs than LDS slots), dense packed input, the batched record form (direct
calls and graph-captured plans), one to three radix passes (slice sizes from
2 to 4096 shards), records that overflow a wave's reservation (atomics), words
already set before the insert, inserts on two streams, inserts into all-zero
words (and each way words change outside an insert), empty and tiny inputs, and
the AUTO path choice.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, make_oracle, random_dna, random_protein

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def seq_for(alphabet, n, seed):
    if alphabet == "protein":
        return random_protein(n, seed, invalid_rate=0.002)
    return random_dna(n, seed, n_rate=0.001, lower_rate=0.01)


def sorted_filter(cfg, bits, slice_bits=0, reserve=0):
    import cusbf_amd as cs

    f = cs.Filter(cfg, bits)
    f.set_insert_path(cs.INSERT_SORTED, slice_bits, reserve)
    return f


def atomic_filter(cfg, bits):
    import cusbf_amd as cs

    f = cs.Filter(cfg, bits)
    f.set_insert_path(cs.INSERT_ATOMIC)
    return f


def oracle_words(k, s, m, h, alphabet, seqs, num_shards):
    o = make_oracle(k, s, m, h, alphabet)
    w = o.new_filter(num_shards)
    for q in seqs:
        if len(q) >= k:
            o.insert(q, w)
    return w


CONFIGS = [
    (31, 28, 16, 4, "dna"),
    (31, 30, 16, 4, "dna"),
    (31, 31, 16, 4, "dna"),
    (31, 28, 21, 4, "dna"),
    (31, 28, 16, 8, "dna"),
    (31, 28, 16, 16, "dna"),  # salted bit addresses
    (31, 28, 24, 4, "dna"),   # minimizer span 8
    (31, 31, 31, 4, "dna"),   # every window its own run: four slot chunks per tile
    (21, 19, 11, 4, "dna"),
    (5, 4, 3, 4, "dna"),      # ~500 runs per tile: two slot chunks
    (3, 2, 2, 4, "dna"),      # ~660 runs per tile: three slot chunks
    (10, 8, 6, 4, "protein"),  # ~330 runs per tile: mostly one chunk of 384 slots
    (12, 10, 6, 4, "protein"),  # the reference example's protein configuration
    (5, 4, 3, 4, "protein"),   # ~500 runs per tile: two chunks of 384 slots
]


@pytest.mark.parametrize("cfg_t", CONFIGS, ids=lambda c: f"{c[4]}-k{c[0]}s{c[1]}m{c[2]}h{c[3]}")
@pytest.mark.parametrize("slice_bits", [12, 5, 1])
def test_sorted_matches_oracle_and_atomic(gpu, cfg_t, slice_bits):
    import cusbf_amd as cs

    k, s, m, h, alphabet = cfg_t
    seq = seq_for(alphabet, 400_003, seed=k * 7 + s + m + h)
    cfg = make_config(k, s, m, h, alphabet)
    bits = 1 << 22  # 2^14 shards: 2 .. 13 slice-index bits, one or two radix passes
    f = sorted_filter(cfg, bits, slice_bits)
    assert f.insert_sequence(seq) == len(seq) - k + 1
    assert f.insert_path().last_path == cs.INSERT_SORTED
    a = atomic_filter(cfg, bits)
    a.insert_sequence(seq)
    assert a.insert_path().last_path == cs.INSERT_ATOMIC
    w = oracle_words(k, s, m, h, alphabet, [seq], f.num_shards())
    assert (a.words() == w).all()
    assert (f.words() == w).all()


def test_sorted_three_passes_and_existing_words(gpu):
    """2^19 shards in slices of 2 -> 18 index bits (8 + 8 + 2); a second insert ORs into
    the words the first left."""
    k, s, m, h = 31, 28, 16, 4
    first = random_dna(600_000, 1, n_rate=0.001)
    second = random_dna(300_000, 2)
    cfg = make_config(k, s, m, h)
    f = sorted_filter(cfg, 1 << 27, slice_bits=1)
    f.insert_sequence(first)
    f.insert_sequence(second)
    w = oracle_words(k, s, m, h, "dna", [first, second], f.num_shards())
    assert (f.words() == w).all()


@pytest.mark.parametrize("reserve", [1, 8])
def test_sorted_region_overflow_falls_back_to_atomics(gpu, reserve):
    """A reservation far below the runs of a tile: most runs take the atomic path."""
    k, s, m, h = 31, 28, 16, 4
    seq = random_dna(500_000, 3, n_rate=0.002)
    cfg = make_config(k, s, m, h)
    f = sorted_filter(cfg, 1 << 24, slice_bits=8, reserve=reserve)
    f.insert_sequence(seq)
    w = oracle_words(k, s, m, h, "dna", [seq], f.num_shards())
    assert (f.words() == w).all()


def test_sorted_dense_packed(gpu):
    import cusbf_amd as cs

    k, s, m, h = 31, 28, 16, 4
    seq = random_dna(300_001, 4)
    cfg = make_config(k, s, m, h)
    words_in = cs.pack_dense_sequence(cfg, seq)
    f = sorted_filter(cfg, 1 << 22, slice_bits=4)
    f.insert_dense_packed(words_in, len(seq))
    assert f.insert_path().last_path == cs.INSERT_SORTED
    w = oracle_words(k, s, m, h, "dna", [seq], f.num_shards())
    assert (f.words() == w).all()


# Generic kernels (no lane strips): DNA triplets, a custom two-symbol alphabet, DNA
# with a findere span past the runtime-span kernels' limit, a protein k past 12 is not
# possible (5-bit symbols), so protein stays on the lane strips.
GENERIC = [
    (3, 2, 2, 4, "triplet"),
    (7, 5, 4, 8, "triplet"),
    (31, 8, 16, 4, "dna"),    # findere span 24
    (31, 28, 4, 16, "dna"),   # minimizer span 28
]


@pytest.mark.parametrize("cfg_t", GENERIC, ids=lambda c: f"{c[4]}-k{c[0]}s{c[1]}m{c[2]}h{c[3]}")
def test_sorted_generic_kernels(gpu, cfg_t):
    """Plain inserts on the generic kernels take the sorted path too: words equal the
    atomic path's and the oracle's, into zero words and into existing ones."""
    import cusbf_amd as cs

    k, s, m, h, alphabet = cfg_t
    width = 3 if alphabet == "triplet" else 1
    seq = random_dna(300_003 * width, k + m, n_rate=0.002)
    more = random_dna(200_001 * width, k + m + 1, n_rate=0.002)
    cfg = make_config(k, s, m, h, alphabet)
    bits = 1 << 22
    f = sorted_filter(cfg, bits)
    assert f.kernel_family() == "generic"
    f.insert_sequence(seq)
    assert f.insert_path().last_path == cs.INSERT_SORTED
    a = atomic_filter(cfg, bits)
    a.insert_sequence(seq)
    w = oracle_words(k, s, m, h, alphabet, [seq], f.num_shards())
    assert (a.words() == w).all()
    assert (f.words() == w).all()
    f.insert_sequence(more)
    a.insert_sequence(more)
    w2 = oracle_words(k, s, m, h, alphabet, [seq, more], f.num_shards())
    assert (f.words() == w2).all() and (a.words() == w2).all()
    g = cs.Filter(cfg, bits)  # AUTO
    g.insert_sequence(seq)
    g.insert_sequence(more)
    assert (g.words() == w2).all()


@pytest.mark.parametrize("cfg_t,want", [((31, 28, 5, 4), "sorted"), ((31, 28, 7, 4), "atomic"),
                                         ((31, 8, 16, 4), "atomic")], ids=["m5", "m7", "m16"])
def test_auto_weighs_hot_shards_on_generic_kernels(gpu, cfg_t, want):
    """AUTO on configurations whose atomic insert uses the generic kernels' hashed tiles.
    A short minimizer (m = 5: 4^5 m-mers, most minimizers among the ~38 with the smallest
    hashes) puts the runs on a few shards whose atomics serialise, so the sorted insert
    (direct tiles) pays; at m = 7 or m = 16 the atomics are not slowed enough to pay for
    direct tiles. The words equal the oracle's either way."""
    import cusbf_amd as cs

    k, s, m, h = cfg_t
    seq = random_dna(20_000_000, k * m + h)
    cfg = make_config(k, s, m, h)
    f = cs.Filter(cfg, 1 << 28)
    assert f.kernel_family() == "generic"
    f.insert_sequence(seq)
    assert f.insert_path().last_path == (cs.INSERT_SORTED if want == "sorted" else cs.INSERT_ATOMIC)
    o = make_oracle(k, s, m, h, "dna")
    w = o.new_filter(f.num_shards())
    o.mt_insert(seq, w)
    assert (f.words() == w).all()


@pytest.mark.parametrize("cfg_t", [(7, 5, 4, 8, "triplet"), (31, 8, 16, 4, "dna")],
                         ids=["triplet-k7", "dna-s8"])
def test_sorted_generic_batched(gpu, cfg_t):
    """Batched inserts (filter.cuh:338-364 record batches, records back to back with
    offsets) on the generic kernels take the sorted path: forced, and under AUTO where the
    cost model prefers it; no window crosses a record start. Words equal the atomic
    batched insert's and the oracle's over the records one by one."""
    import torch

    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    k, s, m, h, alphabet = cfg_t
    width = 3 if alphabet == "triplet" else 1
    rng = np.random.default_rng(k + m)
    lens = rng.integers(0, 400, 30_000)
    lens[:3] = [0, k - 1, k]
    data = random_dna(int(lens.sum()) * width, 77, n_rate=0.002)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    d_data = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    d_off = torch.from_numpy(off).cuda()
    cfg = make_config(k, s, m, h, alphabet)
    o = make_oracle(k, s, m, h, alphabet)
    filters = []
    for path in (cs.INSERT_SORTED, cs.INSERT_ATOMIC, cs.INSERT_AUTO):
        f = cs.Filter(cfg, 1 << 28)
        f.set_insert_path(path)
        bt.insert_batched_async(f, d_data, d_off)
        bt.batched_status(f)
        filters.append(f)
    assert filters[0].insert_path().last_path == cs.INSERT_SORTED
    assert filters[1].insert_path().last_path == cs.INSERT_ATOMIC
    w = o.new_filter(filters[0].num_shards())
    for i in range(lens.size):
        rec = data[off[i] * width:off[i + 1] * width]
        if lens[i] >= k:
            o.insert(rec, w)
    for f in filters:
        assert (f.words() == w).all()


RANGED = [
    (31, 28, 16, 4, "dna"),
    (10, 8, 6, 4, "protein"),
    (31, 23, 14, 4, "dna"),  # runtime-span kernels
]


@pytest.mark.parametrize("cfg_t", RANGED, ids=lambda c: f"{c[4]}-k{c[0]}s{c[1]}m{c[2]}h{c[3]}")
def test_sorted_in_window_ranges(gpu, cfg_t):
    """A workspace limit below one insert's records: the insert runs as several
    sorted inserts over consecutive window ranges (each range after the first ORs
    into the words the earlier ones left) and leaves the oracle's words."""
    import cusbf_amd as cs

    k, s, m, h, alphabet = cfg_t
    seq = seq_for(alphabet, 1_000_003, seed=k + 3 * m)
    cfg = make_config(k, s, m, h, alphabet)
    bits = 1 << 22
    whole = sorted_filter(cfg, bits)
    whole.insert_sequence(seq)
    one = whole.sorted_workspace()
    assert one.last_chunks == 1 and one.limit == 0 and one.allocated > 0
    f = sorted_filter(cfg, bits)
    f.set_sorted_workspace_limit(one.allocated // 4)
    assert f.insert_sequence(seq) == len(seq) - k + 1
    info = f.sorted_workspace()
    assert info.last_chunks >= 4, info
    assert 0 < info.allocated <= one.allocated // 4
    assert f.insert_path().last_path == cs.INSERT_SORTED
    w = oracle_words(k, s, m, h, alphabet, [seq], f.num_shards())
    assert (whole.words() == w).all()
    assert (f.words() == w).all()
    # a second insert of other symbols ORs into the words in ranges again
    more = seq_for(alphabet, 700_001, seed=k + 3 * m + 1)
    f.insert_sequence(more)
    assert f.sorted_workspace().last_chunks >= 3
    w2 = oracle_words(k, s, m, h, alphabet, [seq, more], f.num_shards())
    assert (f.words() == w2).all()


def test_sorted_in_window_ranges_dense(gpu):
    """Dense packed input in window ranges: each range starts on a packed word."""
    import cusbf_amd as cs

    k, s, m, h = 31, 28, 16, 4
    seq = random_dna(1_000_003, 8)
    cfg = make_config(k, s, m, h)
    words_in = cs.pack_dense_sequence(cfg, seq)
    f = sorted_filter(cfg, 1 << 22)
    f.set_sorted_workspace_limit(400_000)
    assert f.insert_dense_packed(words_in, len(seq)) == len(seq) - k + 1
    assert f.sorted_workspace().last_chunks >= 4
    o = make_oracle(k, s, m, h, "dna")
    w = o.new_filter(f.num_shards())
    o.insert_dense(o.pack_dense(seq), len(seq), w)
    assert (f.words() == w).all()


def test_sorted_workspace_limit_too_small(gpu):
    """A limit below one wave tile's workspace: SORTED reports a resource error and
    leaves the words alone; AUTO takes the atomic path. A limit below the workspace
    the filter holds frees it."""
    import cusbf_amd as cs

    k, s, m, h = 31, 28, 16, 4
    seq = random_dna(200_001, 9, n_rate=0.001)
    cfg = make_config(k, s, m, h)
    f = sorted_filter(cfg, 1 << 22)
    f.insert_sequence(seq)
    assert f.sorted_workspace().allocated > 0
    f.clear()
    f.set_sorted_workspace_limit(4096)
    assert f.sorted_workspace().allocated == 0
    with pytest.raises(cs.CusbfError):
        f.insert_sequence(seq)
    assert not f.words().any()
    f.set_insert_path(cs.INSERT_AUTO)
    f.insert_sequence(seq)
    assert f.insert_path().last_path == cs.INSERT_ATOMIC
    assert f.sorted_workspace().last_chunks == 0
    w = oracle_words(k, s, m, h, "dna", [seq], f.num_shards())
    assert (f.words() == w).all()
    f.set_sorted_workspace_limit(0)
    f.set_insert_path(cs.INSERT_SORTED)
    f.insert_sequence(seq)
    assert f.sorted_workspace().last_chunks == 1
    assert (f.words() == w).all()


def test_sorted_record_batch(gpu):
    """Host record batch (records joined by separators, one insert launch)."""
    k, s, m, h = 31, 28, 16, 4
    rng = np.random.default_rng(5)
    recs = [random_dna(int(n), 100 + i) for i, n in enumerate(rng.integers(0, 3000, 400))]
    data = b"".join(recs)
    off = np.cumsum([0] + [len(r) for r in recs])
    ranges = [(int(off[i]), len(r)) for i, r in enumerate(recs)]
    cfg = make_config(k, s, m, h)
    f = sorted_filter(cfg, 1 << 23, slice_bits=6)
    f.insert_record_batch(data, ranges)
    w = oracle_words(k, s, m, h, "dna", recs, f.num_shards())
    assert (f.words() == w).all()


def _batched_inputs(alphabet, n, median, seed):
    from tests.test_gpu_batched import (device, device_offsets, lognormal_lengths, offsets_of,
                                        symbols_for)

    lens = lognormal_lengths(n, median, 1.5, seed=seed)
    lens[:4] = [0, 1, 30, 31]
    off = offsets_of(lens)
    data = symbols_for(alphabet, int(off[-1]), seed=seed + 1)
    return data, off, device(data), device_offsets(off)


@pytest.mark.parametrize("alphabet,k,s,m,claim", [("dna", 31, 28, 16, 0), ("dna", 31, 28, 16, 4),
                                                  ("protein", 10, 8, 6, 0)])
def test_sorted_batched_and_plan(gpu, alphabet, k, s, m, claim):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from tests.test_gpu_batched import oracle_words as batched_oracle_words

    data, off, d_data, d_off = _batched_inputs(alphabet, 3000, 400, seed=k + claim)
    cfg = make_config(k, s, m, 4, alphabet)
    bits = max(1 << 20, int(off[-1]) * 16)
    w = batched_oracle_words(make_oracle(k, s, m, 4, alphabet), data, off, alphabet,
                             cs.Filter(cfg, bits).num_shards())
    f = sorted_filter(cfg, bits, slice_bits=3)
    bt.set_tuning(f, bt.batch_bucket(int(off[-1]), len(off) - 1), bt.INSERT_SORTED,
                  bt.BatchTuning(3, 4, claim))
    bt.insert_batched_async(f, d_data, d_off)
    bt.batched_status(f)
    assert f.insert_path().last_path == cs.INSERT_SORTED
    assert (f.words() == w).all()
    with bt.BatchPlan(f, bt.INSERT, d_data, d_off) as plan:
        for rep in range(3):
            f.clear()
            plan.launch()
            plan.status()
            assert (f.words() == w).all(), rep


def test_sorted_batched_malformed_offsets_leave_filter_clear(gpu):
    from cusbf_amd import batched as bt
    from cusbf_amd.errors import CusbfError
    from tests.test_gpu_batched import device, device_offsets

    f = sorted_filter(make_config(31, 28, 16, 4), 1 << 20, slice_bits=4)
    data = device(random_dna(5000, 9))
    with pytest.raises(CusbfError):
        bt.insert_batched_async(f, data, device_offsets(np.array([0, 3000, 2000, 5000], np.uint64)))
        bt.batched_status(f)
    assert f.load_factor() == 0.0


def test_sorted_two_streams(gpu):
    """Sorted inserts on two streams share the workspace; the second waits for the first."""
    torch = _torch()
    k, s, m, h = 31, 28, 16, 4
    a_seq, b_seq = random_dna(700_000, 11), random_dna(500_000, 12)
    f = sorted_filter(make_config(k, s, m, h), 1 << 24, slice_bits=7)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    d_a = torch.frombuffer(bytearray(a_seq), dtype=torch.uint8).cuda()
    d_b = torch.frombuffer(bytearray(b_seq), dtype=torch.uint8).cuda()
    torch.cuda.synchronize()
    f.insert_sequence_async(d_a, stream=s1)
    f.insert_sequence_async(d_b, stream=s2)
    torch.cuda.synchronize()
    w = oracle_words(k, s, m, h, "dna", [a_seq, b_seq], f.num_shards())
    assert (f.words() == w).all()


def test_fresh_filter_tracking(gpu):
    """An insert into all-zero words (a created or cleared filter) starts each slice from
    zeros instead of reading it. Words changed outside the library's inserts must still
    count: written through the words pointer, refilled by a captured plan replayed after
    a clear, or set by an atomic insert still running on another stream."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from tests.test_gpu_batched import oracle_words as batched_oracle_words

    torch = _torch()
    k, s, m, h = 31, 28, 16, 4
    cfg = make_config(k, s, m, h)
    bits = 1 << 24
    a_seq, b_seq = random_dna(400_000, 21), random_dna(300_000, 22)
    shards = cs.Filter(cfg, bits).num_shards()
    w_a = oracle_words(k, s, m, h, "dna", [a_seq], shards)
    w_b = oracle_words(k, s, m, h, "dna", [b_seq], shards)

    # cleared after an insert: the next insert equals one into a new filter
    f = sorted_filter(cfg, bits, slice_bits=6)
    f.insert_sequence(a_seq)
    f.clear()
    f.insert_sequence(b_seq)
    assert f.insert_path().last_path == cs.INSERT_SORTED
    assert (f.words() == w_b).all()

    # words written through the device pointer before the first insert are kept
    g = sorted_filter(cfg, bits, slice_bits=6)
    t = g.words_tensor()
    rng = np.random.default_rng(5)
    idx = torch.from_numpy(rng.choice(t.numel(), 4096, replace=False)).cuda()
    t[idx] = torch.from_numpy(rng.integers(1, 1 << 62, 4096, dtype=np.int64)).cuda()
    torch.cuda.synchronize()
    planted = g.words()
    assert planted.any()
    g.insert_sequence(b_seq)
    assert g.insert_path().last_path == cs.INSERT_SORTED
    assert (g.words() == (planted | w_b)).all()

    # a captured insert plan replayed after a clear, then an eager sorted insert
    data, off, d_data, d_off = _batched_inputs("dna", 500, 400, seed=23)
    w_plan = batched_oracle_words(make_oracle(k, s, m, h, "dna"), data, off, "dna", shards)
    p = sorted_filter(cfg, bits, slice_bits=6)
    with bt.BatchPlan(p, bt.INSERT, d_data, d_off) as plan:
        p.clear()
        plan.launch()
        plan.status()
        p.insert_sequence(b_seq)
        assert p.insert_path().last_path == cs.INSERT_SORTED
        assert (p.words() == (w_plan | w_b)).all()

    # an atomic insert on one stream, then a sorted insert on another without a wait
    q = cs.Filter(cfg, bits)
    q.set_insert_path(cs.INSERT_ATOMIC)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    d_a = torch.frombuffer(bytearray(a_seq), dtype=torch.uint8).cuda()
    d_b = torch.frombuffer(bytearray(b_seq), dtype=torch.uint8).cuda()
    torch.cuda.synchronize()
    q.insert_sequence_async(d_a, stream=s1)
    q.set_insert_path(cs.INSERT_SORTED, 6, 0)
    q.insert_sequence_async(d_b, stream=s2)
    torch.cuda.synchronize()
    assert q.insert_path().last_path == cs.INSERT_SORTED
    assert (q.words() == (w_a | w_b)).all()


def test_captured_insert_replayed_after_clear(gpu):
    """An insert captured into the caller's graph (inside a capture the atomic path runs)
    resets the filter's zero flag on every replay, so a sorted insert after a replay reads
    the slices; checked on every replay, with eager work between replays."""
    import cusbf_amd as cs

    torch = _torch()
    k, s, m, h = 31, 28, 16, 4
    cfg = make_config(k, s, m, h)
    a_seq, b_seq = random_dna(300_000, 31), random_dna(200_000, 32)
    f = sorted_filter(cfg, 1 << 24, slice_bits=6)
    w_a = oracle_words(k, s, m, h, "dna", [a_seq], f.num_shards())
    w_b = oracle_words(k, s, m, h, "dna", [b_seq], f.num_shards())
    d_a = torch.frombuffer(bytearray(a_seq), dtype=torch.uint8).cuda()
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        f.insert_sequence_async(d_a, stream=side)
    torch.cuda.synchronize()
    for rep in range(3):
        f.clear()
        graph.replay()
        torch.cuda.synchronize()
        assert (f.words() == w_a).all(), rep
        f.insert_sequence(b_seq)
        assert f.insert_path().last_path == cs.INSERT_SORTED
        assert (f.words() == (w_a | w_b)).all(), rep


def test_plan_replay_ordered_with_direct_insert_on_another_stream(gpu):
    """An insert plan that took the sorted path (its apply rewrites whole slices)
    replayed on one stream and a direct sorted insert issued on another stream with
    no host wait between them: the library orders the two, so the words equal the
    oracle's for both inputs (an overlap would lose one side's bits)."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from tests.test_gpu_batched import oracle_words as batched_oracle_words

    torch = _torch()
    k, s, m, h = 31, 28, 16, 4
    cfg = make_config(k, s, m, h)
    bits = 1 << 26
    data, off, d_data, d_off = _batched_inputs("dna", 2000, 1500, seed=41)
    b_seq = random_dna(3_000_000, 42)
    f = sorted_filter(cfg, bits, slice_bits=5)
    shards = f.num_shards()
    w = batched_oracle_words(make_oracle(k, s, m, h, "dna"), data, off, "dna", shards)
    w |= oracle_words(k, s, m, h, "dna", [b_seq], shards)
    d_b = torch.frombuffer(bytearray(b_seq), dtype=torch.uint8).cuda()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with bt.BatchPlan(f, bt.INSERT, d_data, d_off) as plan:
        for rep in range(3):
            f.clear()
            torch.cuda.synchronize()
            plan.launch(s1.cuda_stream)
            f.insert_sequence_async(d_b, stream=s2.cuda_stream)
            torch.cuda.synchronize()
            assert f.insert_path().last_path == cs.INSERT_SORTED
            assert (f.words() == w).all(), rep
            # and the other order: direct insert first, then the replay
            f.clear()
            torch.cuda.synchronize()
            f.insert_sequence_async(d_b, stream=s2.cuda_stream)
            plan.launch(s1.cuda_stream)
            torch.cuda.synchronize()
            assert (f.words() == w).all(), rep


def test_captured_clear_replays(gpu):
    """A clear captured into a graph (fill kernels, no synchronize inside the capture)
    zeroes the words on every replay; an insert after each replay equals the oracle's
    words for that input alone, with eager work between the replays."""
    torch = _torch()
    k, s, m, h = 31, 28, 16, 4
    cfg = make_config(k, s, m, h)
    a_seq, b_seq = random_dna(300_000, 51), random_dna(200_000, 52)
    f = sorted_filter(cfg, 1 << 24, slice_bits=6)
    w_b = oracle_words(k, s, m, h, "dna", [b_seq], f.num_shards())
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        f.clear(side.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        f.insert_sequence(a_seq)
        assert f.words().any()
        graph.replay()
        torch.cuda.synchronize()
        assert not f.words().any(), rep
        f.insert_sequence(b_seq)
        assert (f.words() == w_b).all(), rep


def test_sorted_tiny_and_empty(gpu):
    import cusbf_amd as cs

    cfg = make_config(31, 28, 16, 4)
    f = sorted_filter(cfg, 1 << 20, slice_bits=2)
    assert f.insert_sequence(b"") == 0
    assert f.insert_sequence(b"ACGT" * 7) == 0  # 28 symbols < k
    assert f.load_factor() == 0.0
    one = random_dna(31, 13)
    assert f.insert_sequence(one) == 1
    assert f.insert_path().last_path == cs.INSERT_SORTED
    w = oracle_words(31, 28, 16, 4, "dna", [one], f.num_shards())
    assert (f.words() == w).all()
    allbad = b"N" * 5000
    f.insert_sequence(allbad)
    assert (f.words() == w).all()


def test_auto_path_choice(gpu):
    """AUTO keeps small inserts on the atomic path and moves large ones to the sorted path;
    both give the same words."""
    import cusbf_amd as cs

    torch = _torch()
    cfg = make_config(31, 28, 16, 4)
    f = cs.Filter(cfg, 1 << 24)
    assert f.insert_path().path == cs.INSERT_AUTO
    f.insert_sequence(random_dna(20_000, 14))
    assert f.insert_path().last_path == cs.INSERT_ATOMIC
    big = torch.randint(0, 4, (48_000_000,), dtype=torch.uint8, device="cuda")
    big = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device="cuda")[big.long()]
    g = cs.Filter(cfg, 1 << 24)
    g.insert_sequence_async(big)
    torch.cuda.synchronize()
    assert g.insert_path().last_path == cs.INSERT_SORTED
    a = atomic_filter(cfg, 1 << 24)
    a.insert_sequence_async(big)
    torch.cuda.synchronize()
    assert (g.words() == a.words()).all()


def test_insert_path_arguments(gpu):
    import cusbf_amd as cs
    from cusbf_amd.errors import CusbfError

    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    for bad in [(3, 0, 0), (cs.INSERT_SORTED, 13, 0), (cs.INSERT_SORTED, 0, 1025)]:
        with pytest.raises(CusbfError):
            f.set_insert_path(*bad)
    f.set_insert_path(cs.INSERT_SORTED, 0, 0)
    info = f.insert_path()
    assert (info.path, info.slice_bits, info.reserve_per_tile) == (cs.INSERT_SORTED, 12, 0)
