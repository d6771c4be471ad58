"""GPU parity of the batched record form (records back to back + offsets, one launch).

Expected answers come from the CPU oracle on the same records joined by the
alphabet's separator: a window inside a record must get the oracle's answer
for that window, a window across a record start must be 0, and the filter
words after a batched insert must equal the oracle's words after inserting the
joined records. Covers the lane-strip kernels (DNA and protein front ends),
the generic kernels, every tuning shape, graph-captured plans replayed
several times, and the edge cases of the offsets array.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, make_oracle, random_dna, random_protein

pytestmark = pytest.mark.gpu

SEPARATOR = {"dna": b"N", "protein": b"*", "triplet": b"NNN"}
WIDTH = {"dna": 1, "protein": 1, "triplet": 3}


def _torch():
    import torch

    return torch


def lognormal_lengths(n, median, sigma, seed, lo=0, hi=20000):
    rng = np.random.default_rng(seed)
    lens = np.round(rng.lognormal(np.log(median), sigma, n))
    return np.clip(lens, lo, hi).astype(np.int64)


def symbols_for(alphabet, n_symbols, seed, invalid_rate=0.002) -> bytes:
    if alphabet == "protein":
        return random_protein(n_symbols, seed, invalid_rate=invalid_rate)
    return random_dna(n_symbols * WIDTH[alphabet], seed, n_rate=invalid_rate)


def offsets_of(lens) -> np.ndarray:
    off = np.zeros(len(lens) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    return off


def joined(data: bytes, off: np.ndarray, alphabet: str) -> tuple:
    """Records joined by the separator, and each record's start in the joined text (symbols)."""
    w = WIDTH[alphabet]
    sep = SEPARATOR[alphabet]
    parts, starts, pos = [], [], 0
    for i in range(len(off) - 1):
        a, b = int(off[i]), int(off[i + 1])
        if i:
            parts.append(sep)
            pos += 1
        starts.append(pos)
        parts.append(data[a * w:b * w])
        pos += b - a
    return b"".join(parts), starts


def expected_hits(o, words, data, off, k, alphabet) -> np.ndarray:
    """Batched-layout answers: the oracle's answer for every in-record window, 0 elsewhere."""
    text, starts = joined(data, off, alphabet)
    n_sym = int(off[-1])
    exp = np.zeros(max(0, n_sym - k + 1), dtype=np.uint8)
    if exp.size == 0 or len(text) < k * WIDTH[alphabet]:
        return exp
    hits = o.contains(text, words)
    for i in range(len(off) - 1):
        a, b = int(off[i]), int(off[i + 1])
        n = b - a - k + 1
        if n > 0:
            exp[a:a + n] = hits[starts[i]:starts[i] + n]
    return exp


def oracle_words(o, data, off, alphabet, num_shards):
    words = o.new_filter(num_shards)
    text, _ = joined(data, off, alphabet)
    if len(text):
        o.insert(text, words)
    return words


def device(data: bytes, dtype=None):
    torch = _torch()
    t = torch.frombuffer(bytearray(data) if data else bytearray(1), dtype=torch.uint8)
    return t[:len(data)].cuda()


def device_offsets(off: np.ndarray):
    torch = _torch()
    return torch.from_numpy(off.astype(np.int64)).cuda()


def mutate_every_other(alphabet, data: bytes, off: np.ndarray, seed: int) -> bytes:
    """Same lengths; odd records replaced by fresh random symbols (expected negatives)."""
    w = WIDTH[alphabet]
    out = bytearray(data)
    fresh = symbols_for(alphabet, int(off[-1]), seed)
    for i in range(1, len(off) - 1, 2):
        a, b = int(off[i]) * w, int(off[i + 1]) * w
        out[a:b] = fresh[a:b]
    return bytes(out)


CASES = [
    # (k, s, m, h, alphabet, median length, sigma, records, kernel family)
    (31, 28, 16, 4, "dna", 300, 1.5, 3000, "lane_strip"),
    (31, 30, 16, 4, "dna", 150, 1.0, 2000, "lane_strip"),
    (21, 19, 11, 4, "dna", 200, 1.5, 2000, "lane_strip"),
    (5, 4, 3, 4, "dna", 40, 1.5, 3000, "lane_strip"),
    (3, 2, 2, 4, "dna", 20, 1.5, 3000, "lane_strip"),
    (10, 8, 6, 4, "protein", 300, 1.5, 2000, "lane_strip"),
    (12, 10, 6, 4, "protein", 300, 1.5, 2000, "lane_strip"),
    (5, 4, 3, 4, "protein", 40, 1.5, 3000, "lane_strip"),
    (31, 28, 16, 8, "dna", 300, 1.5, 1000, "lane_strip"),  # plain lane-strip, generic batched
    (17, 15, 9, 4, "dna", 200, 1.5, 1000, "lane_strip_wide"),
    (17, 15, 9, 4, "dna", 200, 1.5, 1000, "generic"),  # forced generic
    (7, 6, 4, 4, "triplet", 60, 1.2, 800, "generic"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[4]}-k{c[0]}s{c[1]}m{c[2]}h{c[3]}")
def test_batched_matches_oracle(gpu, monkeypatch, case):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    k, s, m, h, alphabet, median, sigma, n, family = case
    if family == "generic":
        monkeypatch.setenv("CUSBF_FORCE_GENERIC", "1")
    lens = lognormal_lengths(n, median, sigma, seed=k * 100 + s)
    lens[:5] = [0, 1, k - 1, k, k + 1]  # empty, shorter than k, exactly k
    off = offsets_of(lens)
    data = symbols_for(alphabet, int(off[-1]), seed=7 + k)
    query = mutate_every_other(alphabet, data, off, seed=99 + k)
    cfg = make_config(k, s, m, h, alphabet)
    bits = max(1 << 16, int(off[-1]) * 16)
    f = cs.Filter(cfg, bits)
    assert f.kernel_family() == family

    d_data, d_off, d_query = device(data), device_offsets(off), device(query)
    bt.insert_batched_async(f, d_data, d_off)
    bt.batched_status(f)
    o = make_oracle(k, s, m, h, alphabet)
    w = oracle_words(o, data, off, alphabet, f.num_shards())
    assert (f.words() == w).all()

    n_win = int(off[-1]) - k + 1
    out = torch.full((n_win + 7,), 0xEE, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(f, d_query, d_off, out)
    bt.batched_status(f)
    got = out.cpu().numpy()
    exp = expected_hits(o, w, query, off, k, alphabet)
    assert np.array_equal(got[:n_win], exp)
    assert (got[n_win:] == 0xEE).all()  # nothing written past the windows
    # the mutated odd records are mostly negative, the kept even records all positive
    assert exp.sum() > 0 and exp.sum() < n_win


@pytest.mark.parametrize("alphabet,k", [("dna", 31), ("protein", 10), ("triplet", 7)])
def test_join_records_equals_host_join(gpu, alphabet, k):
    """cusbf_join_records_async: the batched records joined by the separator on the
    device equal the host join (empty and one-symbol records, long records, records
    ending inside and across 16-byte chunks), and the joined buffer through the plain
    record-batch query answers like the batched query."""
    from cusbf_amd import batched as bt

    torch = _torch()
    lens = lognormal_lengths(3000, 120, 1.8, seed=k, hi=5000)
    lens[:8] = [0, 0, 1, 2, 15, 16, 17, 0]
    off = offsets_of(lens)
    data = symbols_for(alphabet, int(off[-1]), seed=k + 11)
    cfg = make_config(k, k - 2, k // 2, 4, alphabet)
    import cusbf_amd as cs

    f = cs.Filter(cfg, 1 << 22)
    want, _ = joined(data, off, alphabet)
    got = bt.join_records_async(f, device(data), device_offsets(off))
    torch.cuda.synchronize()
    assert got.numel() == len(want)
    assert bytes(got.cpu().numpy().tobytes()) == want
    with pytest.raises(cs.CusbfError):
        bt.join_records_async(f, device(data), device_offsets(off), d_out=got[: len(want) - 1])
    # the joined layout is what the batched query reads without separators
    bt.insert_batched_async(f, device(data), device_offsets(off))
    bt.batched_status(f)
    n_win = int(off[-1]) - k + 1
    out = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(f, device(data), device_offsets(off), out)
    bt.batched_status(f)
    hits = f.contains_sequence(want)
    _, starts = joined(data, off, alphabet)
    got_b = out.cpu().numpy()
    for i in range(len(lens)):
        n = int(lens[i]) - k + 1
        if n > 0:
            a = int(off[i])
            assert np.array_equal(got_b[a:a + n], hits[starts[i]:starts[i] + n])


def test_batched_equals_plain_record_batch_words(gpu):
    """Batched insert == the plain record-batch path (host records joined by separators)."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    lens = lognormal_lengths(2500, 1000, 1.5, seed=3, lo=31, hi=200000)
    off = offsets_of(lens)
    data = symbols_for("dna", int(off[-1]), seed=5)
    cfg = make_config(31, 28, 16, 4)
    a = cs.Filter(cfg, int(off[-1]) * 22)
    b = cs.Filter(cfg, int(off[-1]) * 22)
    bt.insert_batched_async(a, device(data), device_offsets(off))
    bt.batched_status(a)
    ranges = [(int(off[i]), int(lens[i])) for i in range(len(lens))]
    b.insert_record_batch(data, ranges)
    assert (a.words() == b.words()).all()


TUNINGS = [(3, 1, 0), (3, 4, 1), (3, 4, 2), (3, 2, 8), (3, 8, 64), (3, 3, 4), (3, 1, 1), (3, 32, 0),
           (3, 16, 16), (4, 4, 2)]


@pytest.mark.parametrize("alphabet,k,s,m", [("dna", 31, 28, 16), ("protein", 10, 8, 6)])
def test_batched_tunings_identical(gpu, alphabet, k, s, m):
    """Every launch shape (variant, persistent blocks, claim size) gives the oracle's words
    and answers, on repeated launches (work-queue kernels with several blocks per compute
    unit once returned different answers from run to run)."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    lens = lognormal_lengths(6000, 150, 1.5, seed=11)
    off = offsets_of(lens)
    data = symbols_for(alphabet, int(off[-1]), seed=12)
    query = mutate_every_other(alphabet, data, off, seed=13)
    cfg = make_config(k, s, m, 4, alphabet)
    d_data, d_off, d_query = device(data), device_offsets(off), device(query)
    n_win = int(off[-1]) - k + 1
    o = make_oracle(k, s, m, 4, alphabet)
    bits = int(off[-1]) * 16
    want_words = oracle_words(o, data, off, alphabet, cs.Filter(cfg, bits).num_shards())
    want_hits = expected_hits(o, want_words, query, off, k, alphabet)
    for wps, bpc, claim in TUNINGS:
        if wps not in bt.batch_variants(cs.Filter(cfg, 1 << 12)):
            continue  # no kernel variant for this register budget
        f = cs.Filter(cfg, bits)
        t = bt.BatchTuning(wps, bpc, claim)
        for op in (bt.INSERT, bt.INSERT_SORTED, bt.CONTAINS):
            bt.set_tuning(f, bt.batch_bucket(int(off[-1]), len(lens)), op, t)
        bt.insert_batched_async(f, d_data, d_off)
        bt.batched_status(f)
        assert (f.words() == want_words).all(), (wps, bpc, claim)
        for rep in range(3):
            out = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
            bt.contains_batched_async(f, d_query, d_off, out)
            bt.batched_status(f)
            assert np.array_equal(out.cpu().numpy(), want_hits), (wps, bpc, claim, rep)


def test_batched_protein_four_waves_rejected(gpu):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from cusbf_amd.errors import CusbfError

    f = cs.Filter(make_config(10, 8, 6, 4, "protein"), 1 << 20)
    with pytest.raises(CusbfError) as e:
        bt.set_tuning(f, 0, bt.INSERT, bt.BatchTuning(4, 4, 1))
    assert e.value.category.value == 3
    with pytest.raises(CusbfError):
        bt.set_tuning(f, 0, bt.INSERT, bt.BatchTuning(3, 0, 1))
    with pytest.raises(CusbfError):
        bt.set_tuning(f, 0, bt.INSERT, bt.BatchTuning(3, 4, 65))
    with pytest.raises(CusbfError):
        bt.set_tuning(f, 4, bt.INSERT, bt.BatchTuning(3, 4, 1))
    bt.set_tuning(f, 2, bt.CONTAINS, bt.BatchTuning(3, 5, 0))
    assert bt.get_tuning(f, 2, bt.CONTAINS) == bt.BatchTuning(3, 5, 0)


def test_batched_plan_replays(gpu):
    """A graph-captured plan: every replay equals the eager call, with eager work between."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    lens = lognormal_lengths(4000, 500, 1.5, seed=21, lo=1)
    off = offsets_of(lens)
    data = symbols_for("dna", int(off[-1]), seed=22)
    query = mutate_every_other("dna", data, off, seed=23)
    cfg = make_config(31, 28, 16, 4)
    d_data, d_off, d_query = device(data), device_offsets(off), device(query)
    n_win = int(off[-1]) - 30

    eager = cs.Filter(cfg, int(off[-1]) * 22)
    bt.insert_batched_async(eager, d_data, d_off)
    e_out = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(eager, d_query, d_off, e_out)
    bt.batched_status(eager)
    e_words, e_hits = eager.words(), e_out.cpu().numpy()

    f = cs.Filter(cfg, int(off[-1]) * 22)
    out = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    with bt.BatchPlan(f, bt.INSERT, d_data, d_off) as ins, \
            bt.BatchPlan(f, bt.CONTAINS, d_query, d_off, out) as qry:
        for rep in range(3):
            f.clear()
            ins.launch()
            ins.status()
            assert (f.words() == e_words).all(), rep
            out.fill_(0x55)  # eager work between replays
            torch.cuda.synchronize()
            qry.launch()
            qry.status()
            assert np.array_equal(out.cpu().numpy(), e_hits), rep
        ins.launch()  # re-insert is idempotent
        ins.status()
        assert (f.words() == e_words).all()


@pytest.mark.parametrize("lens", [
    [992], [991, 1, 992], [993, 992, 31, 30], [1024] * 40, [31] * 500, [0, 0, 0, 50, 0],
    [31], [5000, 0, 0, 1, 2, 3, 4000], [30] * 100, [20000],
], ids=lambda l: f"n{len(l)}-sum{sum(l)}")
def test_batched_tile_seams_and_degenerate_records(gpu, lens):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    lens = np.array(lens, dtype=np.int64)
    off = offsets_of(lens)
    data = symbols_for("dna", int(off[-1]), seed=int(lens.sum()) % 1000, invalid_rate=0.0)
    cfg = make_config(31, 28, 16, 4)
    f = cs.Filter(cfg, 1 << 22)
    d_data, d_off = device(data), device_offsets(off)
    bt.insert_batched_async(f, d_data, d_off)
    bt.batched_status(f)
    o = make_oracle(31, 28, 16, 4)
    w = oracle_words(o, data, off, "dna", f.num_shards())
    assert (f.words() == w).all()
    n_win = int(off[-1]) - 30
    if n_win <= 0:
        return
    out = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(f, d_data, d_off, out)
    bt.batched_status(f)
    assert np.array_equal(out.cpu().numpy(), expected_hits(o, w, data, off, 31, "dna"))


def test_batched_empty_and_short_inputs(gpu):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    torch = _torch()
    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    empty = torch.zeros(1, dtype=torch.uint8, device="cuda")
    bt.insert_batched_async(f, empty, device_offsets(np.zeros(1, np.uint64)), num_symbols=0)
    bt.batched_status(f)
    short = device(b"ACGT" * 5)
    bt.insert_batched_async(f, short, device_offsets(np.array([0, 20], np.uint64)))
    bt.contains_batched_async(f, short, device_offsets(np.array([0, 20], np.uint64)), empty,
                              output_size=0)
    bt.batched_status(f)
    assert f.load_factor() == 0.0


def test_batched_malformed_offsets(gpu):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from cusbf_amd.errors import CusbfError

    torch = _torch()
    data = device(random_dna(5000, 1))
    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    out = torch.full((5000,), 7, dtype=torch.uint8, device="cuda")
    for bad in ([0, 3000, 2000, 5000], [1, 2000, 5000], [0, 2000, 4999], [0, 2000, 6000]):
        bt.insert_batched_async(f, data, device_offsets(np.array(bad, np.uint64)))
        with pytest.raises(CusbfError) as e:
            bt.batched_status(f)
        assert e.value.category.value == 3
        assert "non-decreasing" in str(e.value)
        bt.contains_batched_async(f, data, device_offsets(np.array(bad, np.uint64)), out)
        with pytest.raises(CusbfError):
            bt.batched_status(f)
    assert f.load_factor() == 0.0  # rejected launches did no work
    assert (out.cpu().numpy() == 7).all()
    # a good call afterwards reports success
    bt.insert_batched_async(f, data, device_offsets(np.array([0, 2500, 5000], np.uint64)))
    bt.batched_status(f)
    assert f.load_factor() > 0.0
    with pytest.raises(CusbfError) as e:
        bt.contains_batched_async(f, data, device_offsets(np.array([0, 5000], np.uint64)), out,
                                  output_size=100)
    assert "too small" in str(e.value)


def test_words_tensor_views_device_words(gpu):
    import cusbf_amd as cs

    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    f.insert_sequence(random_dna(20000, 4))
    t = f.words_tensor()
    assert t.numel() == f.num_shards() * 4
    assert np.array_equal(t.cpu().numpy().view(np.uint64), f.words())


def test_tuner_picks_a_measured_candidate(gpu):
    import cusbf_amd as cs
    from cusbf_amd import batched as bt, tune

    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 24)
    res = tune.tune_bucket(f, 0, sample_symbols=1 << 21, reps=1)
    for op in ("insert", "insert_sorted", "contains"):
        r = res[op]
        assert len(r.timings) == len(tune.candidates(f)) == 20
        assert r.best_ms == min(t[3] for t in r.timings)
    table = tune.tune(f, buckets=(1,), sample_symbols=1 << 21, reps=1)
    for name, op in (("insert", bt.INSERT), ("insert_sorted", bt.INSERT_SORTED),
                     ("contains", bt.CONTAINS)):
        assert bt.get_tuning(f, 1, op) == bt.BatchTuning(**table["bucket1"][name]["best"])
    assert f.load_factor() == 0.0  # tuning never writes the target filter
    assert f.insert_path().path == cs.INSERT_AUTO
    g = cs.Filter(make_config(7, 6, 4, 4, "triplet"), 1 << 20)  # generic kernels: nothing to tune
    assert tune.candidates(g) == [] and tune.tune(g) == {}


def test_per_record_launch_loop_matches_batched(gpu):
    """The plain per-record baseline loop answers exactly like the batched form."""
    import ctypes

    import cusbf_amd as cs
    from cusbf_amd import _lib, batched as bt
    from cusbf_amd import workloads as W

    torch = _torch()
    lens = lognormal_lengths(3000, 200, 1.5, seed=41, lo=0)
    off = offsets_of(lens)
    data = symbols_for("protein", int(off[-1]), seed=42)
    cfg = make_config(10, 8, 6, 4, "protein")
    d_data, d_off = device(data), device_offsets(off)
    n_win = int(off[-1]) - 9
    a = cs.Filter(cfg, 1 << 24)
    b = cs.Filter(cfg, 1 << 24)
    bt.insert_batched_async(a, d_data, d_off)
    out_a = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    bt.contains_batched_async(a, d_data, d_off, out_a)
    bt.batched_status(a)
    lib, synth = _lib.load(), W.synth_lib()
    fn_i = ctypes.cast(lib.cusbf_insert_sequence_async, ctypes.c_void_p).value
    fn_q = ctypes.cast(lib.cusbf_contains_sequence_async, ctypes.c_void_p).value
    h_off = np.ascontiguousarray(off)
    out_b = torch.zeros(n_win, dtype=torch.uint8, device="cuda")
    assert synth.cusbf_bench_each_record(fn_i, 0, b.handle, d_data.data_ptr(), 1, h_off.ctypes.data,
                                         len(lens), 10, None, 0) == 0
    assert synth.cusbf_bench_each_record(fn_q, 1, b.handle, d_data.data_ptr(), 1, h_off.ctypes.data,
                                         len(lens), 10, out_b.data_ptr(), 0) == 0
    torch.cuda.synchronize()
    assert torch.equal(a.words_tensor(), b.words_tensor())
    assert torch.equal(out_a, out_b)


@pytest.mark.parametrize("alphabet", ["dna", "protein"])
def test_batched_full_size_words_equal_oracle(gpu, alphabet):
    """The measured batched workload at full size: 1M seeded log-normal records (~3.08 G
    symbols) inserted by a graph plan (sorted path) into a 2^36-bit filter; every filter
    word equals the multithreaded oracle's over the records joined by separators."""
    import cusbf_amd as cs
    from cusbf_amd import batched as bt
    from cusbf_amd import workloads as W

    torch = _torch()
    torch.cuda.empty_cache()
    if alphabet == "dna":
        k, s, m, letters, sep, seed = 31, 28, 16, W.DNA_LETTERS, ord("N"), 31
        cfg = cs.Config(k, s, m, 4)
    else:
        k, s, m, letters, sep, seed = 10, 8, 6, W.PROTEIN_LETTERS, ord("*"), 37
        cfg = cs.Config(k, s, m, 4, alphabet=cs.ProteinAlphabet)
    data, d_off, lens = W.generate_batched(W.BatchedSpec(records=1_000_000, letters=letters, seed=seed))
    kmers = int(np.maximum(lens.astype(np.int64) - k + 1, 0).sum())
    f = cs.Filter(cfg, int(21.4 * kmers))
    assert f.num_shards() == 1 << 28
    f.set_insert_path(cs.INSERT_SORTED)
    with bt.BatchPlan(f, bt.INSERT, data, d_off) as plan:
        plan.launch()
        plan.status()
    host = W.join_records(data, lens, sep).cpu().numpy()
    del data, d_off
    o = make_oracle(k, s, m, 4, alphabet)
    w = o.new_filter(f.num_shards())
    o.mt_insert(host, w)
    assert np.array_equal(f.words(), w)
