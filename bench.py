"""Benchmark: Super Bloom insert & query k-mers/s on one MI355X, with HBM roofline.

One step = clear the filter, insert every k-mer of the synthetic genome
(one record batch, chromosomes joined by separators), then query every k-mer
window of the read set (150 bp reads joined by separators). Inputs are
resident in HBM before timing starts; HIP events on the launch stream bracket
the insert and query kernels (the clear is outside the event pair, as the
reference clears outside its timed region).

value = (valid inserted k-mers + valid queried k-mers) / (insert + query kernel time)

Prints one JSON line. --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "insert & query k-mers/s at 1 GPU (% HBM roofline) + p50/p99 serve latency"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
INSERT_PATHS = {"auto": 0, "atomic": 1, "sorted": 2}  # CUSBF_INSERT_*


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def guarded(name: str, fn, *args, **kwargs):
    """Runs an auxiliary leg: a failure is logged and recorded in the JSON line as
    {"error": ...} instead of ending the run (the headline figures do not depend on it)."""
    try:
        return fn(*args, **kwargs)
    except Exception as e:  # noqa: BLE001 - any failure of a side leg is reported, not raised
        log(f"[bench] {name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="human", choices=["human", "nematode"])
    ap.add_argument("--reads", type=int, default=0, help="override the number of reads")
    ap.add_argument("--insert-path", default="auto", choices=["auto", "atomic", "sorted"],
                    help="insert path of the lane-strip kernels (auto: sorted when it pays)")
    ap.add_argument("--nematode", default="on", choices=["on", "off"],
                    help="nematode-scale leg (100 Mbp, 85.7 bits/k-mer, 20M reads) and the "
                         "reference's FPR densities on its genome")
    ap.add_argument("--fpr-kmers", type=int, default=1_000_000_000,
                    help="uniform random query k-mers of the FPR densities sweep")
    ap.add_argument("--placement-check", type=int, default=0,
                    help="after the timed steps, time the query on this many further filters "
                         "of the same size, each allocated while the earlier ones are kept")
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--probes", default="on", choices=["on", "off"],
                    help="measure stream-copy / random-sector ceilings after the timed steps")
    ap.add_argument("--serve", default="on", choices=["on", "off"],
                    help="serving scenario: p50/p99 latency under Poisson arrivals")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-qps", type=float, default=100.0)
    ap.add_argument("--dense", default="on", choices=["on", "off"],
                    help="dense packed input beside byte input (device and host forms)")
    ap.add_argument("--dense-host-bases", type=int, default=400_000_000)
    ap.add_argument("--fpr", default="on", choices=["on", "off"],
                    help="empirical FPR on the foreign reads at s = 28, 30, 31 (m = 16) and m = 21")
    ap.add_argument("--fastx", default="on", choices=["on", "off"],
                    help="FASTA insert / FASTQ query through files (text ingest throughput)")
    ap.add_argument("--fastx-bases", type=int, default=400_000_000)
    ap.add_argument("--fastx-reads", type=int, default=4_000_000)
    ap.add_argument("--batched", default="on", choices=["on", "off"],
                    help="batched uneven-length form (DNA and protein) vs plain launches")
    ap.add_argument("--batched-records", type=int, default=1_000_000)
    ap.add_argument("--batched-reps", type=int, default=5)
    ap.add_argument("--batched-baselines", default="on", choices=["on", "off"],
                    help="time the per-record and host record-batch baselines (1M launches)")
    ap.add_argument("--tuning-in", default="",
                    help="apply a saved sweep (written by --tuning-out) instead of tuning")
    ap.add_argument("--tuning-out", default="",
                    help="write the autotuner's full sweep (JSON) to this path")
    return ap.parse_args()


def committed_traffic(workload: str, kernel: str):
    """Calibrated HBM bytes per launch of `kernel` from the committed counter summary."""
    name = {"insert": "insert_kernel<31, 28, 16, 4, Dna2, false, 3>",
            "query": "contains_kernel<31, 28, 16, 4, Dna2, false, 3>"}[kernel]
    path = os.path.join(ROOT, "profiles", f"dna_insert_query_{workload}_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.split("|")]
        if len(cells) >= 5 and cells[1] == name:
            return float(cells[4])
    return None


# Published FPR of the reference at 21.4 bits/k-mer (percent; its benchmark
# instantiates m = 21, k = 31, h = 4), for context beside the measured values.
PUBLISHED_FPR_M21 = {"s28": 0.848, "s30": 0.951, "s31": 1.593}


def fpr_scenario(genome, reads, wl, out, stream_handle) -> dict:
    """Empirical false-positive rate: positives among the windows of the foreign
    (uniform random) reads, which are reads [genome_frac * n, n) of the query
    buffer, over their valid windows. s = 28 comes from the timed query's
    output; s = 30, 31 (m = 16) and s = 28 with m = 21 insert the genome into a
    filter of their own and query the foreign reads."""
    import torch

    import cusbf_amd as cs

    k, rl = wl.k, wl.reads.read_len
    n = wl.reads.n_reads
    g = int(n * wl.reads.genome_frac)
    start = g * (rl + 1)
    foreign = reads[start:]
    n_win = foreign.numel() - k + 1
    negatives = (n - g) * (rl - k + 1)  # foreign reads hold no invalid symbol
    genome_windows = genome.numel() - k + 1
    res = {"negative_kmers": negatives, "foreign_reads": n - g,
           "bits_per_kmer_requested": wl.bits_per_kmer, "m16": {}, "m21": {},
           "published_m21_percent": PUBLISHED_FPR_M21}
    pos = int(cs.count_positive(out[start:start + n_win]))
    res["m16"]["s28"] = pos / negatives
    for s_, m_ in [(30, 16), (31, 16), (28, 21)]:
        f = cs.Filter(cs.Config(k, s_, m_, wl.h), int(wl.bits_per_kmer * genome_windows))
        f.insert_sequence_async(genome, stream=stream_handle)
        o = torch.empty(n_win, dtype=torch.uint8, device="cuda")
        f.contains_sequence_async(foreign, o, stream=stream_handle)
        pos = int(cs.count_positive(o))
        res[f"m{m_}"][f"s{s_}"] = pos / negatives
        res["bits_per_kmer_effective"] = f.filter_bits() / genome_windows
        f.close()
        del o
        torch.cuda.empty_cache()
    log(f"[bench] FPR m=16 s=28/30/31: " + " / ".join(f"{100 * res['m16'][f's{x}']:.3f}%" for x in (28, 30, 31))
        + f"; m=21 s=28: {100 * res['m21']['s28']:.3f}%")
    return res


# The reference's published C. elegans throughput (README.md:37-44, figure (a)),
# G k-mers/s, and its FPR at 21.4 / 85.7 / 342.6 bits per k-mer (README.md:59-63),
# percent, with m = 21 (benchmarks/fpr_fastx_sweep.cu:295).
PUBLISHED_NEMATODE = {"rtx_pro_6000": {"insert": 64.0, "query": 140.0},
                      "gh200": {"insert": 37.0, "query": 53.0}}
PUBLISHED_FPR_DENSITIES = {"21.4": {"s28": 0.848, "s30": 0.951, "s31": 1.593},
                           "85.7": {"s28": 0.091, "s30": 0.107, "s31": 0.210},
                           "342.6": {"s28": 0.0095, "s30": 0.0114, "s31": 0.0264}}


def timed_insert_query(filt, genome, reads, out, sh, steps: int, warmup: int):
    """Median insert (filter cleared outside the timed region) and query seconds."""
    import torch

    stream = torch.cuda.current_stream()
    ti, tq = [], []
    for i in range(warmup + steps):
        filt.clear(sh)
        ev = _events(3)
        ev[0].record(stream)
        filt.insert_sequence_async(genome, stream=sh)
        ev[1].record(stream)
        filt.contains_sequence_async(reads, out, stream=sh)
        ev[2].record(stream)
        ev[2].synchronize()
        if i >= warmup:
            ti.append(ev[0].elapsed_time(ev[1]) / 1e3)
            tq.append(ev[1].elapsed_time(ev[2]) / 1e3)
    return statistics.median(ti), statistics.median(tq)


def nematode_scenario(steps: int, warmup: int, fpr_kmers: int) -> dict:
    """The nematode-scale plain path: 100 Mbp genome (35% GC, 6 chromosomes, 5%
    repeats) at 85.7 bits per k-mer, 20 M 150 bp reads; insert and query timed like
    the headline workload. Then the reference's FPR protocol
    (benchmarks/fpr_fastx_sweep.cu): the genome inserted into filters of 2^31,
    2^33 and 2^35 bits (21.4 / 85.7 / 342.6 bits per k-mer), `fpr_kmers` uniform
    random k-mers queried, s = 28 / 30 / 31 with m = 21 (the reference's
    benchmark) and m = 16."""
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    wl = W.WORKLOADS["nematode"]
    genome, _ = W.generate_genome(wl.genome)
    reads = W.generate_reads(wl.reads, genome)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    windows = genome.numel() - wl.k + 1
    filt = cs.Filter(cs.Config(wl.k, wl.s, wl.m, wl.h), int(wl.bits_per_kmer * windows))
    out = torch.empty(reads.numel() - wl.k + 1, dtype=torch.uint8, device="cuda")
    v_ins, s_ins = filt.count_windows_runs(genome)
    v_q, s_q = filt.count_windows_runs(reads)
    ti, tq = timed_insert_query(filt, genome, reads, out, sh, steps, warmup)
    positives = int(cs.count_positive(out))
    bytes_ins = genome.numel() + 64 * s_ins
    bytes_q = reads.numel() + out.numel() + 32 * s_q
    res = {"workload": {"genome_bp": wl.genome.total_len, "chromosomes": wl.genome.n_chrom,
                        "gc": wl.genome.gc, "repeat_frac": wl.genome.repeat_frac,
                        "reads": wl.reads.n_reads, "read_len": wl.reads.read_len,
                        "k": wl.k, "s": wl.s, "m": wl.m, "hash_count": wl.h,
                        "bits_per_kmer_requested": wl.bits_per_kmer,
                        "filter_bits": filt.filter_bits(),
                        "bits_per_kmer_effective": filt.filter_bits() / max(1, v_ins)},
           "insert_path": {1: "atomic", 2: "sorted"}[filt.insert_path().last_path],
           "insert_ms": ti * 1e3, "query_ms": tq * 1e3,
           "insert_kmers_per_s": v_ins / ti, "query_kmers_per_s": v_q / tq,
           "kmers_per_s": (v_ins + v_q) / (ti + tq),
           "insert_valid_kmers": v_ins, "insert_shard_runs": s_ins,
           "query_valid_kmers": v_q, "query_shard_runs": s_q,
           "insert_roofline_frac": bytes_ins / ti / HBM_PEAK,
           "query_roofline_frac": bytes_q / tq / HBM_PEAK,
           "query_positive_frac": positives / max(1, v_q),
           "published_reference_G_kmers_per_s": PUBLISHED_NEMATODE}
    log(f"[nematode] insert {ti * 1e3:.3f} ms ({v_ins / ti / 1e9:.1f} G k-mers/s), query "
        f"{tq * 1e3:.3f} ms ({v_q / tq / 1e9:.1f} G k-mers/s)")
    filt.close()
    # the sizing of the reference's published figures, bit_ceil(16 x k-mers) bits
    # (benchmarks/benchmark_common.cuh:178-186): 2^31 bits for this genome
    ref_bits = 1 << (16 * windows - 1).bit_length()
    g = cs.Filter(cs.Config(wl.k, wl.s, wl.m, wl.h), ref_bits)
    ti2, tq2 = timed_insert_query(g, genome, reads, out, sh, steps, warmup)
    res["reference_sizing"] = {"filter_bits": g.filter_bits(),
                               "bits_per_kmer_effective": g.filter_bits() / max(1, v_ins),
                               "insert_ms": ti2 * 1e3, "query_ms": tq2 * 1e3,
                               "insert_kmers_per_s": v_ins / ti2, "query_kmers_per_s": v_q / tq2,
                               "insert_path": {1: "atomic", 2: "sorted"}[g.insert_path().last_path]}
    log(f"[nematode] at 2^{ref_bits.bit_length() - 1} bits (the reference's sizing): insert "
        f"{v_ins / ti2 / 1e9:.1f}, query {v_q / tq2 / 1e9:.1f} G k-mers/s")
    g.close()
    del reads, out
    torch.cuda.empty_cache()

    q = W.uniform_dna(fpr_kmers + wl.k - 1, seed=0xDEADBEEF)
    hits = torch.empty(fpr_kmers, dtype=torch.uint8, device="cuda")
    fpr = {"query_kmers": fpr_kmers, "insert_kmers": windows,
           "published_m21_percent": PUBLISHED_FPR_DENSITIES, "m21": {}, "m16": {}}
    for log2_bits, label in ((31, "21.4"), (33, "85.7"), (35, "342.6")):
        for m_ in (21, 16):
            row = {}
            for s_ in (28, 30, 31):
                f = cs.Filter(cs.Config(wl.k, s_, m_, wl.h), 1 << log2_bits)
                f.insert_sequence_async(genome, stream=sh)
                f.contains_sequence_async(q, hits, stream=sh)
                row[f"s{s_}"] = 100.0 * int(cs.count_positive(hits)) / fpr_kmers
                f.close()
            row["bits_per_kmer"] = (1 << log2_bits) / windows
            fpr[f"m{m_}"][label] = row
    log("[nematode] FPR m=21 (percent): " + "; ".join(
        f"{b}: " + " / ".join(f"{fpr['m21'][b][f's{x}']:.4f}" for x in (28, 30, 31))
        for b in ("21.4", "85.7", "342.6")))
    res["fpr_percent"] = fpr
    res["param_sweep"] = param_sweep_points(genome, q[:genome.numel()], hits, sh)
    del q, hits, genome
    torch.cuda.empty_cache()
    return res


# Points of the reference's parameter grids (benchmarks/param_sweep.cu:263-591): DNA k = 31,
# s = 16..31, m = 8..31, h = 4..16, corners and middle, and protein k = 12, s = 6..12,
# m = 4..12. (31, 28, 16, 8), (31, 28, 16, 16), (31, 28, 24, 4), (31, 30, 20, 8) and
# (31, 31, 31, 4) have lane strips compiled for them; the others run the runtime-span lane
# strips. Each point is also timed on the generic kernels (one window per thread,
# CUSBF_FORCE_GENERIC), which served them before the runtime-span kernels existed.
SWEEP_POINTS = [(31, 28, 16, 8), (31, 16, 8, 4), (31, 20, 12, 8), (31, 24, 16, 12), (31, 18, 10, 16),
                (31, 22, 18, 4), (31, 26, 12, 8), (31, 18, 26, 8), (31, 24, 29, 4), (31, 20, 31, 8),
                (31, 28, 16, 16), (31, 28, 24, 4), (31, 30, 20, 8), (31, 31, 31, 4)]
PROTEIN_SWEEP_POINTS = [(12, 6, 4, 16), (12, 8, 6, 8), (12, 9, 8, 4), (12, 10, 6, 4), (12, 12, 12, 4)]


def _sweep_filter(cfg, bits, generic: bool):
    import cusbf_amd as cs

    if generic:
        os.environ["CUSBF_FORCE_GENERIC"] = "1"
    try:
        return cs.Filter(cfg, bits)
    finally:
        os.environ.pop("CUSBF_FORCE_GENERIC", None)


def _sweep_times(f, seq, query, hits, sh, reps):
    import torch

    stream = torch.cuda.current_stream()
    n_q = query.numel() - f.config.k + 1
    ti, tq = [], []
    for _ in range(reps):
        f.clear(sh)
        ev = _events(3)
        ev[0].record(stream)
        f.insert_sequence_async(seq, stream=sh)
        ev[1].record(stream)
        f.contains_sequence_async(query, hits[:n_q], stream=sh)
        ev[2].record(stream)
        ev[2].synchronize()
        ti.append(ev[0].elapsed_time(ev[1]) / 1e3)
        tq.append(ev[1].elapsed_time(ev[2]) / 1e3)
    return statistics.median(ti), statistics.median(tq)


def param_sweep_points(genome, query, hits, sh, reps: int = 3) -> list:
    """Insert the nematode genome and query a uniform sequence of its length at each DNA
    sweep point (a uniform 20-letter protein sequence of the same length, inserted and
    queried with a second one, at the protein points), filter at 21.4 bits per k-mer;
    medians of `reps` HIP-event timings, on the dispatched kernels and on the generic ones."""
    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    out = []
    protein = W.uniform_letters(genome.numel(), b"ACDEFGHIKLMNPQRSTVWY", seed=41)
    protein_q = W.uniform_letters(genome.numel(), b"ACDEFGHIKLMNPQRSTVWY", seed=42)
    points = [("dna", p) for p in SWEEP_POINTS] + [("protein", p) for p in PROTEIN_SWEEP_POINTS]
    for alphabet, (k, s_, m_, h_) in points:
        if alphabet == "dna":
            cfg, seq, qry = cs.Config(k, s_, m_, h_), genome, query
        else:
            cfg, seq, qry = cs.Config(k, s_, m_, h_, alphabet=cs.ProteinAlphabet), protein, protein_q
        n_ins = seq.numel() - k + 1
        n_q = qry.numel() - k + 1
        f = _sweep_filter(cfg, int(21.4 * n_ins), generic=False)
        ti, tq = _sweep_times(f, seq, qry, hits, sh, reps)
        row = {"alphabet": alphabet, "k": k, "s": s_, "m": m_, "h": h_, "kernels": f.kernel_family(),
               "insert_path": {1: "atomic", 2: "sorted"}[f.insert_path().last_path],
               "insert_kmers_per_s": n_ins / ti, "query_kmers_per_s": n_q / tq}
        f.close()
        g = _sweep_filter(cfg, int(21.4 * n_ins), generic=True)
        gi, gq = _sweep_times(g, seq, qry, hits, sh, reps)
        g.close()
        row["generic_insert_kmers_per_s"] = n_ins / gi
        row["generic_query_kmers_per_s"] = n_q / gq
        out.append(row)
        log(f"[sweep] {alphabet} k={k} s={s_} m={m_} h={h_} {row['kernels']}: insert "
            f"{row['insert_kmers_per_s'] / 1e9:.1f}, query {row['query_kmers_per_s'] / 1e9:.1f} G k-mers/s "
            f"(generic {row['generic_insert_kmers_per_s'] / 1e9:.1f} / {row['generic_query_kmers_per_s'] / 1e9:.1f})")
    del protein, protein_q
    return out


def dense_scenario(genome, wl, stream_handle, host_bases: int, reps: int = 3) -> dict:
    """Dense packed wire format beside byte input (dense_packed_throughput.cu's
    fixtures) on a seeded uniform ACGT sequence of the genome's length (dense
    words hold no invalid symbol, so the chromosome separators of the genome
    would read differently): packed on the device (2 bits per base,
    LSB-first), then byte vs dense insert and query -- the query is the
    inserted sequence, every k-mer present -- timed with HIP events (filter
    cleared outside the timed region, median of `reps`); and the host forms
    (host bytes / host words in, one synchronous call each, transfers
    included) on its first `host_bases` bases."""
    import numpy as np
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    cfg = cs.Config(wl.k, wl.s, wl.m, wl.h)
    genome = W.uniform_dna(genome.numel(), seed=31)
    n = genome.numel()
    windows = n - wl.k + 1
    words = torch.empty((n + 31) // 32, dtype=torch.int64, device="cuda")
    f = cs.Filter(cfg, int(wl.bits_per_kmer * windows))
    out = torch.empty(windows, dtype=torch.uint8, device="cuda")

    def timed(op, setup=None):
        ts = []
        for _ in range(reps):
            if setup:
                setup()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            op()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b) / 1e3)
        return statistics.median(ts)

    clear = lambda: f.clear(stream_handle)
    t_pack = timed(lambda: cs.pack_dense_sequence_device(cfg, genome, n, words, stream=stream_handle))
    t_bi = timed(lambda: f.insert_sequence_async(genome, stream=stream_handle), clear)
    t_bq = timed(lambda: f.contains_sequence_async(genome, out, stream=stream_handle))
    byte_pos = int(cs.count_positive(out))
    t_di = timed(lambda: f.insert_dense_packed_async(words, n, stream=stream_handle), clear)
    t_dq = timed(lambda: f.contains_dense_packed_async(words, n, out, stream=stream_handle))
    dense_pos = int(cs.count_positive(out))
    res = {"symbols": n, "kmers": windows, "packed_bytes": words.numel() * 8,
           "pack_device_GBps": n / t_pack / 1e9,
           "device": {"byte_insert_kmers_per_s": windows / t_bi, "dense_insert_kmers_per_s": windows / t_di,
                      "byte_query_kmers_per_s": windows / t_bq, "dense_query_kmers_per_s": windows / t_dq,
                      "byte_insert_ms": t_bi * 1e3, "dense_insert_ms": t_di * 1e3,
                      "byte_query_ms": t_bq * 1e3, "dense_query_ms": t_dq * 1e3},
           "all_positive": byte_pos == windows and dense_pos == windows}
    # host forms on a sample: host bytes in / host words in
    hn = min(n, host_bases)
    hw = hn - wl.k + 1
    host_bytes = genome[:hn].cpu().numpy()
    host_words = words[: (hn + 31) // 32].cpu().numpy().view(np.uint64).copy()
    if hn % 32:  # the sample's last word: only its first hn % 32 bases
        host_words[-1] &= np.uint64((1 << (2 * (hn % 32))) - 1)
    g = cs.Filter(cfg, int(wl.bits_per_kmer * hw))

    def wall(op):
        ts = []
        for _ in range(reps):
            g.clear()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            op()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    def wall_query(op):
        ts, ans = [], None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ans = op()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts), ans

    t_hbi = wall(lambda: g.insert_sequence(host_bytes))
    t_hdi = wall(lambda: g.insert_dense_packed(host_words, hn))
    g.insert_sequence(host_bytes)
    t_hbq, hb = wall_query(lambda: g.contains_sequence(host_bytes))
    t_hdq, hd = wall_query(lambda: g.contains_dense_packed(host_words, hn))
    # cusbf_host_overhead.cu: host bytes in, answers into device memory, against
    # the device forms on the same bases (HIP events)
    d_out = torch.empty(hw, dtype=torch.uint8, device="cuda")
    t_hbqd, _ = wall_query(lambda: g.contains_sequence(host_bytes, d_output=d_out))
    hd_dev_equal = bool(np.array_equal(d_out.cpu().numpy(), hb))
    sample = genome[:hn]
    t_dbi = timed(lambda: g.insert_sequence_async(sample, stream=stream_handle),
                  lambda: g.clear(stream_handle))
    t_dbq = timed(lambda: g.contains_sequence_async(sample, d_out, stream=stream_handle))
    res["host"] = {"bases": hn, "byte_insert_kmers_per_s": hw / t_hbi, "dense_insert_kmers_per_s": hw / t_hdi,
                   "byte_query_kmers_per_s": hw / t_hbq, "dense_query_kmers_per_s": hw / t_hdq,
                   "byte_query_to_device_kmers_per_s": hw / t_hbqd,
                   "byte_insert_ms": t_hbi * 1e3, "byte_query_ms": t_hbq * 1e3,
                   "byte_query_to_device_ms": t_hbqd * 1e3,
                   "device_insert_ms": t_dbi * 1e3, "device_query_ms": t_dbq * 1e3,
                   "answers_equal": bool(np.array_equal(hb, hd)) and hd_dev_equal}
    f.close()
    g.close()
    del words, out, genome, d_out, sample
    torch.cuda.empty_cache()
    d = res["device"]
    log(f"[bench] dense: device insert byte {d['byte_insert_ms']:.2f} ms / dense {d['dense_insert_ms']:.2f} ms, "
        f"query byte {d['byte_query_ms']:.2f} ms / dense {d['dense_query_ms']:.2f} ms; host insert "
        f"{res['host']['byte_insert_kmers_per_s'] / 1e9:.2f} / {res['host']['dense_insert_kmers_per_s'] / 1e9:.2f} G k-mers/s; "
        f"host forms on {hn / 1e6:.0f} Mbp: insert {res['host']['byte_insert_ms']:.1f} ms (device "
        f"{res['host']['device_insert_ms']:.1f}), query to host {res['host']['byte_query_ms']:.1f} ms, to device "
        f"{res['host']['byte_query_to_device_ms']:.1f} ms (device {res['host']['device_query_ms']:.1f})")
    return res


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(genome, reads, wl, threads: int):
    """CPU Super Bloom (the oracle's multithreaded sliding form) on a bounded sample."""
    import numpy as np

    from oracle import oracle as orc

    threads = threads or orc.default_threads()
    ins_bytes = min(genome.numel(), 400_000_000)
    read_stride = wl.reads.read_len + 1
    q_reads = min(wl.reads.n_reads, 4_000_000)
    g = genome[:ins_bytes].cpu().numpy()
    r = reads[: q_reads * read_stride - 1].cpu().numpy()
    o = orc.Oracle(wl.k, wl.s, wl.m, wl.h)
    n_ins = o.kmer_count(g.size)
    shards = o.num_shards(int(wl.bits_per_kmer * n_ins))
    words = o.new_filter(shards)
    t0 = time.perf_counter()
    o.mt_insert(g, words, threads)
    t1 = time.perf_counter()
    hits = o.mt_contains(r, words, threads)
    t2 = time.perf_counter()
    v_ins, _ = o.count(g[: min(g.size, 20_000_000)], shards)  # valid fraction estimate
    frac = v_ins / max(1, o.kmer_count(min(g.size, 20_000_000)))
    ins_valid = n_ins * frac
    q_valid = q_reads * (wl.reads.read_len - wl.k + 1)
    ins_rate = ins_valid / (t1 - t0)
    q_rate = q_valid / (t2 - t1)
    return {
        "value": (ins_valid + q_valid) / ((t1 - t0) + (t2 - t1)),
        "unit": "k-mers/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "hardware_concurrency": orc.hardware_threads(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "sample": (f"insert first {ins_bytes / 1e6:.0f} Mbp of the genome, query first "
                   f"{q_reads / 1e6:.0f}M reads; {threads} threads; "
                   f"insert {ins_rate / 1e6:.0f} M/s, query {q_rate / 1e6:.0f} M/s; "
                   f"{int(hits.sum())} positives"),
        "insert_kmers_per_s": ins_rate,
        "query_kmers_per_s": q_rate,
        "seconds": (t2 - t0),
    }


def serving_scenario(filt, genome, reads, wl, n_requests: int, qps: float):
    """Serving form on the resident filter vs plain per-request launches.

    Requests: Poisson arrivals at `qps`; sizes log-uniform over 1..1M reads of
    150 bp (long-read requests get the same byte sizes, at least one read);
    half short reads (1% substitutions), half 10-50 kbp long reads (5% mixed
    substitutions and indels). The server admits requests into 4 batch slots
    under a 16 GB staging budget. Latency = arrival to answers ready.
    """
    import threading

    import numpy as np
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    rng = np.random.default_rng(4242)
    stride = wl.reads.read_len + 1
    pool_reads = min(wl.reads.n_reads, 8_000_000)
    short = reads[: pool_reads * stride - 1].cpu().numpy()
    d_long, long_off = W.generate_long_reads(genome, 20_000, seed=77)
    long_pool = d_long.cpu().numpy()
    del d_long
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_requests))
    reqs = W.serving_requests(short, wl.reads.read_len, long_pool, long_off, n_requests, seed=4243)
    total_bytes = sum(int(r[0].size) for r in reqs)
    # answers land in one pre-faulted arena reused by every run
    from cusbf_amd.serve import normalized_size, ranges_array

    win = [filt.record_kmer_count(normalized_size(*ranges_array(rg), 1)) for _, rg in reqs]
    starts = np.concatenate([[0], np.cumsum(win)]).astype(np.int64)
    arena = np.ones(int(starts[-1]) + 1, dtype=np.uint8)
    outs = [arena[starts[i]:starts[i] + max(1, win[i])] for i in range(n_requests)]

    def pct(x, q):
        return float(np.percentile(np.asarray(x), q)) * 1e3

    def run_server(burst: bool, arr=None):
        arr = arrivals if arr is None else arr
        lat = [0.0] * n_requests
        kmers = [0] * n_requests
        positives = [0] * n_requests
        with cs.Server(filt, staging_bytes=16 << 30, slots=4) as srv:
            pend = [None] * n_requests
            lag = [0.0] * n_requests
            t0 = time.perf_counter()

            def submitter():
                for i, (seq, rg) in enumerate(reqs):
                    if not burst:
                        dt = t0 + arr[i] - time.perf_counter()
                        if dt > 0:
                            time.sleep(dt)
                        lag[i] = max(0.0, time.perf_counter() - (t0 + arr[i]))
                    pend[i] = srv.submit(seq, rg, out=outs[i])
                    ready[i].set()

            ready = [threading.Event() for _ in range(n_requests)]
            th = threading.Thread(target=submitter)
            th.start()
            for i in range(n_requests):
                ready[i].wait()
                _, rep, l = srv.wait(pend[i])
                lat[i] = l + lag[i]
                kmers[i] = rep.queried_kmers
                positives[i] = rep.positive_kmers
                pend[i] = None
            th.join()
            span = time.perf_counter() - t0
            st = srv.stats()
        return lat, kmers, positives, span, st

    def run_plain(burst: bool):
        lat = [0.0] * n_requests
        kmers = [0] * n_requests
        positives = [0] * n_requests
        t0 = time.perf_counter()
        for i, (seq, rg) in enumerate(reqs):
            if not burst:
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            _, rep = cs.query_record_batch_hits(filt, seq, rg, out=outs[i])
            lat[i] = time.perf_counter() - (t0 + (0.0 if burst else arrivals[i]))
            kmers[i] = rep.queried_kmers
            positives[i] = rep.positive_kmers
        return lat, kmers, positives, time.perf_counter() - t0

    log(f"[serve] {n_requests} requests, {total_bytes / 1e9:.2f} GB, Poisson {qps:.0f} QPS")
    s_lat, s_km, s_pos, s_span, s_st = run_server(False)
    p_lat, p_km, p_pos, p_span = run_plain(False)
    log(f"[serve] server p50 {pct(s_lat, 50):.1f} ms p99 {pct(s_lat, 99):.1f} ms; "
        f"plain p50 {pct(p_lat, 50):.1f} ms p99 {pct(p_lat, 99):.1f} ms")
    b_lat, b_km, b_pos, b_span, b_st = run_server(True)
    pb_lat, pb_km, pb_pos, pb_span = run_plain(True)
    log(f"[serve] burst: server {sum(b_km) / b_span / 1e9:.2f} G k-mers/s, plain "
        f"{sum(pb_km) / pb_span / 1e9:.2f} G k-mers/s; span {b_span:.2f} s: fill "
        f"{b_st['fill_seconds']:.2f} s, launch {b_st['launch_seconds']:.2f} s, device "
        f"{b_st['device_seconds']:.2f} s, answer copies {b_st['copy_seconds']:.2f} s")
    # load sweep: Poisson arrivals at fractions of the saturation rate (the burst's
    # requests per second)
    sat_qps = n_requests / b_span
    unit = np.cumsum(np.random.default_rng(4244).exponential(1.0, n_requests))
    sweep = []
    for frac in (0.25, 0.5, 0.8):
        q_lat, q_km, q_pos, q_span, _ = run_server(False, unit / (frac * sat_qps))
        identical_q = q_pos == b_pos
        sweep.append({"fraction_of_saturation": frac, "qps": frac * sat_qps,
                      "p50_ms": pct(q_lat, 50), "p99_ms": pct(q_lat, 99),
                      "kmers_per_s": sum(q_km) / q_span, "answers_identical": bool(identical_q)})
        log(f"[serve] {frac:.2f} x saturation ({frac * sat_qps:.0f} QPS): p50 {pct(q_lat, 50):.1f} ms "
            f"p99 {pct(q_lat, 99):.1f} ms")
    identical = (s_pos == p_pos == b_pos == pb_pos) and (s_km == p_km)
    return {
        "requests": n_requests, "qps": qps, "bytes": total_bytes,
        "server_p50_ms": pct(s_lat, 50), "server_p99_ms": pct(s_lat, 99),
        "plain_p50_ms": pct(p_lat, 50), "plain_p99_ms": pct(p_lat, 99),
        "server_kmers_per_s": sum(s_km) / s_span, "plain_kmers_per_s": sum(p_km) / p_span,
        "burst_server_kmers_per_s": sum(b_km) / b_span,
        "burst_plain_kmers_per_s": sum(pb_km) / pb_span,
        "burst_speedup": (sum(b_km) / b_span) / max(1e-9, sum(pb_km) / pb_span),
        "server_batches": s_st["batches"], "server_max_requests_per_batch": s_st["max_requests_per_batch"],
        "burst_span_s": b_span,
        "burst_stage_seconds": {k: b_st[k] for k in ("fill_seconds", "launch_seconds",
                                                     "device_seconds", "copy_seconds")},
        "answers_identical": bool(identical),
        "saturation_qps": sat_qps,
        "load_sweep": sweep,
        "staging_bytes": 16 << 30, "slots": 4,
    }


def _events(n):
    import torch

    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def _equal_chunked(a, b, step=1 << 28) -> bool:
    import torch

    if a.numel() != b.numel():
        return False
    return all(torch.equal(a[i:i + step], b[i:i + step]) for i in range(0, a.numel(), step))


def fastx_scenario(genome, reads, wl, n_bases: int, n_reads: int) -> dict:
    """FASTX ingest: a FASTA of the genome's first `n_bases` (records of 50 Mbp,
    60 bases per line) inserted from a file, then a FASTQ of the first `n_reads`
    150 bp reads (CRLF lines) queried from a file, on a fresh filter. Times the
    whole call (mmap, parse, stage, kernels, reports); beside it the parse-only
    time of the same text on the host, and a check that the filter words equal
    the record-batch path's."""
    import ctypes
    import shutil
    import tempfile

    import numpy as np

    import cusbf_amd as cs
    from cusbf_amd import _lib as cs_lib
    from cusbf_amd.errors import check

    n_bases = min(n_bases, genome.numel())
    g = genome[:n_bases].cpu().numpy()
    rec = 50_000_000
    parts = []
    for i, a in enumerate(range(0, n_bases, rec)):
        chunk = g[a:a + rec]
        pad = (-chunk.size) % 60
        rows = np.concatenate([chunk, np.full(pad, 0, np.uint8)]).reshape(-1, 60)
        body = np.concatenate([rows, np.full((rows.shape[0], 1), 10, np.uint8)], axis=1).tobytes()
        if pad:
            body = body[:-(pad + 1)] + b"\n"
        parts.append(b">chunk%d synthetic\n" % i + body)
    fasta = b"".join(parts)
    stride = wl.reads.read_len + 1
    n_reads = min(n_reads, (reads.numel() + 1) // stride)
    r = reads[: n_reads * stride - 1].cpu().numpy()
    r = np.concatenate([r, np.zeros(1, np.uint8)]).reshape(n_reads, stride)[:, :-1]
    qual = np.full_like(r, ord("I"))

    def cols(text: bytes):
        return np.tile(np.frombuffer(text, np.uint8), (n_reads, 1))

    idx = np.arange(n_reads, dtype=np.int64)
    digits = np.stack([(idx // 10 ** j) % 10 for j in range(9, -1, -1)], axis=1).astype(np.uint8) + 48
    fastq = np.concatenate([cols(b"@r"), digits, cols(b"\r\n"), r, cols(b"\r\n+\r\n"), qual,
                            cols(b"\r\n")], axis=1).tobytes()
    tmp = tempfile.mkdtemp(prefix="cusbf_fastx_")
    try:
        fa_path, fq_path = os.path.join(tmp, "genome.fa"), os.path.join(tmp, "reads.fq")
        with open(fa_path, "wb") as fh:
            fh.write(fasta)
        with open(fq_path, "wb") as fh:
            fh.write(fastq)
        cfg = cs.Config(wl.k, wl.s, wl.m, wl.h)
        filt = cs.Filter(cfg, int(wl.bits_per_kmer * n_bases))
        t0 = time.perf_counter()
        ins = filt.insert_fastx_file(fa_path)
        t_ins = time.perf_counter() - t0
        t0 = time.perf_counter()
        q = filt.query_fastx_file(fq_path)
        t_q = time.perf_counter() - t0
        # parse alone, no device work and no consumer: FASTA split by host threads,
        # FASTQ parsed line by line
        lib = cs_lib.load()
        n_rec = ctypes.c_uint64()
        t0 = time.perf_counter()
        none = cs_lib.FASTX_CONSUMER()  # no consumer
        check(lib, lib.cusbf_parse_fastx_buffer_chunked(fasta, len(fasta), b"genome.fa", 0, 0, none, None,
                                                        ctypes.byref(n_rec)))
        t_parse_fa = time.perf_counter() - t0
        t0 = time.perf_counter()
        check(lib, lib.cusbf_parse_fastx_buffer_chunked(fastq, len(fastq), b"reads.fq", 0, 0, none, None,
                                                        ctypes.byref(n_rec)))
        t_parse_fq = time.perf_counter() - t0
        # the record-batch path on the same records gives the same words
        ref = cs.Filter(cfg, filt.filter_bits())
        seqs = [p.split(b"\n", 1)[1].replace(b"\n", b"") for p in parts]
        off = np.cumsum([0] + [len(x) for x in seqs])
        ref.insert_record_batch(b"".join(seqs), [(int(off[i]), len(x)) for i, x in enumerate(seqs)])
        same = bool(np.array_equal(ref.words(), filt.words()))
        ref.close()
        filt.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    log(f"[fastx] FASTA {len(fasta) / 1e9:.2f} GB insert {t_ins:.2f} s, "
        f"FASTQ {len(fastq) / 1e9:.2f} GB query {t_q:.2f} s")
    return {
        "fasta_bytes": len(fasta), "fastq_bytes": len(fastq),
        "insert_seconds": t_ins, "query_seconds": t_q,
        "insert_text_GBps": len(fasta) / t_ins / 1e9, "query_text_GBps": len(fastq) / t_q / 1e9,
        "insert_kmers_per_s": ins.inserted_kmers / t_ins, "query_kmers_per_s": q.queried_kmers / t_q,
        "insert_report": [ins.records_indexed, ins.indexed_bases, ins.inserted_kmers],
        "query_report": [q.records_queried, q.queried_bases, q.queried_kmers, q.positive_kmers],
        "host_parse_only_fasta_GBps": len(fasta) / t_parse_fa / 1e9,
        "host_parse_only_fastq_GBps": len(fastq) / t_parse_fq / 1e9,
        "words_equal_record_batch_path": same,
    }


def batched_scenario(alphabet: str, n_records: int, reps: int, tuning_out: str = "",
                     baselines: bool = True, tuning_in: str = "") -> dict:
    """Batched uneven-length form: 1M log-normal records (median 1 kbp, sigma 1.5,
    clipped to [31 bp, 2 Mbp]) inserted and queried by graph-captured persistent
    launches, against (a) one plain launch per record, (b) the plain host record
    batch path (normalise, stage, one launch) and (c) one plain launch over the
    records joined by separators on the device. Answers and filter words are
    compared in full."""
    import ctypes

    import numpy as np
    import torch

    import cusbf_amd as cs
    from cusbf_amd import _lib, batched as bt, tune
    from cusbf_amd import workloads as W
    from cusbf_amd.serve import ranges_array

    if alphabet == "dna":
        cfg = cs.Config(31, 28, 16, 4)
        letters, sep = W.DNA_LETTERS, ord("N")
    else:
        cfg = cs.Config(10, 8, 6, 4, alphabet=cs.ProteinAlphabet)
        letters, sep = W.PROTEIN_LETTERS, ord("*")
    k = cfg.k
    spec = W.BatchedSpec(records=n_records, letters=letters, seed=31 if alphabet == "dna" else 37)
    data, d_off, lens = W.generate_batched(spec)
    query = W.mutate_odd_records(data, lens, letters, spec.seed + 1)
    total = data.numel()
    record_kmers = int(np.maximum(lens - k + 1, 0).sum())
    log(f"[batched:{alphabet}] {lens.size} records, {total / 1e9:.2f} G symbols, "
        f"{record_kmers / 1e9:.2f} G in-record k-mers")
    filt = cs.Filter(cfg, int(21.4 * record_kmers))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    t0 = time.perf_counter()
    if tuning_in:
        table = tune.load(tuning_in)
        tune.apply(filt, table)
    else:
        table = tune.tune(filt, log=log)
    tune_s = time.perf_counter() - t0
    if tuning_out:
        tune.save(table, tuning_out)
    bucket = bt.batch_bucket(total, lens.size)
    v_ins, s_ins = bt.count_batched_windows_runs(filt, data, d_off)
    v_q, s_q = bt.count_batched_windows_runs(filt, query, d_off)

    windows = total - k + 1
    out_b = torch.empty(windows, dtype=torch.uint8, device="cuda")
    ins = bt.BatchPlan(filt, bt.INSERT, data, d_off)
    qry = bt.BatchPlan(filt, bt.CONTAINS, query, d_off, out_b)
    reps = max(1, reps)
    ti, tq = [], []
    for r in range(reps + 1):
        filt.clear(sh)
        ev = _events(3)
        ev[0].record(stream)
        ins.launch(sh)
        ev[1].record(stream)
        qry.launch(sh)
        ev[2].record(stream)
        ev[2].synchronize()
        if r:  # first pass is warm-up
            ti.append(ev[0].elapsed_time(ev[1]) / 1e3)
            tq.append(ev[1].elapsed_time(ev[2]) / 1e3)
    ins.status(sh)
    qry.status(sh)
    bi, bq = statistics.median(ti), statistics.median(tq)
    positives = int(cs.count_positive(out_b))
    words_b = filt.words_tensor()
    log(f"[batched:{alphabet}] graph launches: insert {bi * 1e3:.2f} ms, query {bq * 1e3:.2f} ms "
        f"({v_ins / bi / 1e9:.1f} / {v_q / bq / 1e9:.1f} G k-mers/s)")

    # (c) plain single launch over the records joined by separators on the device
    # (cusbf_join_records_async; checked against a torch join of the same records)
    rec_idx = W.record_index(lens)
    norm_a = bt.join_records_async(filt, data, d_off)
    norm_q = bt.join_records_async(filt, query, d_off)
    js = []
    for _ in range(3):
        ev = _events(2)
        ev[0].record(stream)
        bt.join_records_async(filt, data, d_off, d_out=norm_a, stream=sh)
        bt.join_records_async(filt, query, d_off, d_out=norm_q, stream=sh)
        ev[1].record(stream)
        ev[1].synchronize()
        js.append(ev[0].elapsed_time(ev[1]) / 1e3 / 2)
    join_s = statistics.median(js)
    join_checked = torch.equal(norm_q, W.join_records(query, lens, sep, rec_idx))
    f2 = cs.Filter(cfg, filt.filter_bits())
    out_n = torch.empty(norm_q.numel() - k + 1, dtype=torch.uint8, device="cuda")
    ni, nq = [], []
    for r in range(reps + 1):
        f2.clear(sh)
        ev = _events(3)
        ev[0].record(stream)
        f2.insert_sequence_async(norm_a, stream=sh)
        ev[1].record(stream)
        f2.contains_sequence_async(norm_q, out_n, stream=sh)
        ev[2].record(stream)
        ev[2].synchronize()
        if r:
            ni.append(ev[0].elapsed_time(ev[1]) / 1e3)
            nq.append(ev[1].elapsed_time(ev[2]) / 1e3)
    pi_, pq_ = statistics.median(ni), statistics.median(nq)
    words_equal_norm = torch.equal(f2.words_tensor(), words_b)
    same_norm = True
    step = 1 << 28
    for a in range(0, windows, step):
        b = min(windows, a + step)
        idx = torch.arange(a, b, device="cuda", dtype=torch.int64) + rec_idx[a:b]
        same_norm = same_norm and torch.equal(out_n[idx], out_b[a:b])
    del norm_a, norm_q, out_n, rec_idx
    f2.close()
    torch.cuda.empty_cache()

    lib = _lib.load()
    kmers = v_ins + v_q
    batched_rate = kmers / (bi + bq)
    result = {
        "workload": {"records": int(lens.size), "symbols": int(total), "median": spec.median,
                     "sigma": spec.sigma, "clip": [spec.min_len, spec.max_len],
                     "alphabet": alphabet, "k": cfg.k, "s": cfg.s, "m": cfg.m,
                     "hash_count": cfg.hash_count, "filter_bits": filt.filter_bits(),
                     "bits_per_kmer_requested": 21.4, "length_bucket": bucket,
                     "insert_valid_kmers": v_ins, "insert_shard_runs": s_ins,
                     "query_valid_kmers": v_q, "query_shard_runs": s_q},
        "kernels": filt.kernel_family(),
        "tuning": {key: {op: r["best"] | {"ms": r["best_ms"]} for op, r in ops.items()}
                   for key, ops in table.items()},
        "tuning_seconds": tune_s,
        "batched_insert_ms": bi * 1e3,
        "batched_query_ms": bq * 1e3,
        "batched_insert_kmers_per_s": v_ins / bi,
        "batched_query_kmers_per_s": v_q / bq,
        "batched_kmers_per_s": batched_rate,
        "batched_insert_roofline_frac": (total * cfg.symbol_width + 64 * s_ins) / bi / HBM_PEAK,
        "batched_query_roofline_frac": (total * cfg.symbol_width + windows + 32 * s_q) / bq / HBM_PEAK,
        "joined_plain_insert_ms": pi_ * 1e3,
        "joined_plain_query_ms": pq_ * 1e3,
        "joined_plain_kmers_per_s": kmers / (pi_ + pq_),
        "join_ms_per_input": join_s * 1e3,
        "speedup_vs_joined_plain_kernels": batched_rate / (kmers / (pi_ + pq_)),
        "speedup_vs_joined_plain_incl_join": batched_rate / (kmers / (pi_ + pq_ + 2 * join_s)),
        "query_positive_kmers": positives,
        "identical": {"words_joined_plain": bool(words_equal_norm),
                      "answers_joined_plain": bool(same_norm),
                      "device_join_equals_torch_join": bool(join_checked)},
    }
    if not baselines:
        return result

    # (a) one plain launch per record, issued from native code
    synth = W.synth_lib()
    f3 = cs.Filter(cfg, filt.filter_bits())
    out_r = torch.zeros(windows, dtype=torch.uint8, device="cuda")
    h_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64))
    fn_ins = ctypes.cast(lib.cusbf_insert_sequence_async, ctypes.c_void_p).value
    fn_q = ctypes.cast(lib.cusbf_contains_sequence_async, ctypes.c_void_p).value
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = synth.cusbf_bench_each_record(fn_ins, 0, f3.handle, data.data_ptr(), cfg.symbol_width,
                                       h_off.ctypes.data, lens.size, k, None, sh)
    torch.cuda.synchronize()
    ri = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc |= synth.cusbf_bench_each_record(fn_q, 1, f3.handle, query.data_ptr(), cfg.symbol_width,
                                        h_off.ctypes.data, lens.size, k, out_r.data_ptr(), sh)
    torch.cuda.synchronize()
    rq = time.perf_counter() - t0
    if rc:
        raise RuntimeError(f"per-record launches failed with status {rc}")
    words_equal_rec = torch.equal(f3.words_tensor(), words_b)
    same_rec = _equal_chunked(out_r, out_b)
    del out_r
    f3.close()
    log(f"[batched:{alphabet}] per-record launches: insert {ri:.2f} s, query {rq:.2f} s")

    # (b) plain host record batch: host records -> normalise -> stage -> one launch
    h_data = data.cpu().numpy()
    h_query = query.cpu().numpy()
    rr = np.stack([h_off[:-1] * cfg.symbol_width, lens.astype(np.uint64) * cfg.symbol_width], 1)
    arr, n = ranges_array(rr)
    f4 = cs.Filter(cfg, filt.filter_bits())
    rep_i = _lib.CInsertReport()
    rep_q = _lib.CQueryReport()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = lib.cusbf_insert_record_batch(f4.handle, h_data.ctypes.data, h_data.size, arr, n, sh,
                                       ctypes.byref(rep_i))
    hi = time.perf_counter() - t0
    t0 = time.perf_counter()
    st |= lib.cusbf_query_record_batch(f4.handle, h_query.ctypes.data, h_query.size, arr, n, sh,
                                       ctypes.byref(rep_q))
    hq = time.perf_counter() - t0
    if st:
        raise RuntimeError(f"host record batch failed with status {st}")
    words_equal_host = torch.equal(f4.words_tensor(), words_b)
    f4.close()
    log(f"[batched:{alphabet}] host record batch: insert {hi:.2f} s, query {hq:.2f} s")

    result.update({
        "per_record_insert_s": ri,
        "per_record_query_s": rq,
        "per_record_kmers_per_s": kmers / (ri + rq),
        "host_record_batch_insert_s": hi,
        "host_record_batch_query_s": hq,
        "host_record_batch_kmers_per_s": kmers / (hi + hq),
        "speedup_vs_per_record": batched_rate / (kmers / (ri + rq)),
        "speedup_vs_host_record_batch": batched_rate / (kmers / (hi + hq)),
        "host_record_batch_positive_kmers": int(rep_q.positive_kmers),
    })
    result["identical"].update({
        "words_per_record": bool(words_equal_rec),
        "answers_per_record": bool(same_rec),
        "words_host_record_batch": bool(words_equal_host),
        "positives_host_record_batch": int(rep_q.positive_kmers) == positives})
    return result


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    import cusbf_amd as cs
    from cusbf_amd import workloads as W

    if not torch.cuda.is_available():
        log("no GPU visible")
        return 1
    nematode = None
    if args.nematode == "on" and args.workload == "human":
        log("[bench] nematode-scale leg")
        nematode = guarded("nematode", nematode_scenario, args.steps, args.warmup, args.fpr_kmers)
    wl = W.WORKLOADS[args.workload]
    if args.reads:
        wl.reads.n_reads = args.reads
    log(f"[bench] workload {wl.name}: generating genome ({wl.genome.total_len / 1e9:.2f} Gbp)")
    t0 = time.perf_counter()
    genome, chrom_lens = W.generate_genome(wl.genome)
    log(f"[bench] generating {wl.reads.n_reads / 1e6:.0f}M reads")
    reads = W.generate_reads(wl.reads, genome)
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s "
        f"(genome {genome.numel() / 1e9:.2f} GB, reads {reads.numel() / 1e9:.2f} GB)")

    cfg = cs.Config(wl.k, wl.s, wl.m, wl.h)
    genome_windows = genome.numel() - wl.k + 1
    filt = cs.Filter(cfg, int(wl.bits_per_kmer * genome_windows))
    filt.set_insert_path(INSERT_PATHS[args.insert_path])
    out = torch.empty(reads.numel() - wl.k + 1, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    # exact work counts for the byte model (valid windows, shard runs)
    v_ins, s_ins = filt.count_windows_runs(genome)
    v_q, s_q = filt.count_windows_runs(reads)
    log(f"[bench] insert: {v_ins} valid k-mers, {s_ins} shard runs; "
        f"query: {v_q} valid k-mers, {s_q} shard runs; kernels={filt.kernel_family()}")

    def step(record=None):
        filt.clear(sh)
        if record:
            record[0].record(stream)
        filt.insert_sequence_async(genome, stream=sh)
        if record:
            record[1].record(stream)
        filt.contains_sequence_async(reads, out, stream=sh)
        if record:
            record[2].record(stream)

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps)")

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    for i in range(args.steps):
        step(events[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    t_ins = [e[0].elapsed_time(e[1]) / 1e3 for e in events]
    t_q = [e[1].elapsed_time(e[2]) / 1e3 for e in events]
    ti, tq = statistics.median(t_ins), statistics.median(t_q)
    log(f"[bench] timed {args.steps} steps in {wall:.3f}s; insert median {ti * 1e3:.2f} ms, "
        f"query median {tq * 1e3:.2f} ms")

    # insert into a populated filter: the genome again into the filter holding it
    # (the words are not known to be zero, so the sorted apply reads every slice)
    pop = []
    for _ in range(3):
        ev = _events(2)
        ev[0].record(stream)
        filt.insert_sequence_async(genome, stream=sh)
        ev[1].record(stream)
        ev[1].synchronize()
        pop.append(ev[0].elapsed_time(ev[1]) / 1e3)
    t_pop = statistics.median(pop)
    log(f"[bench] insert into the populated filter: {t_pop * 1e3:.2f} ms")

    # the insert in window ranges: a sorted workspace limit of a quarter of the single
    # insert's, so the genome runs as several sorted inserts one after another (each
    # range after the first loads and stores the slices it touches); the words must
    # equal the single insert's
    ranged = None
    if filt.insert_path().last_path == cs.INSERT_SORTED:
        def measure_ranges() -> dict:
            whole_ws = filt.sorted_workspace().allocated
            # host copies of the words: a device view of them would hand out the words
            # pointer, after which clears no longer mark the words as zero
            ref_words = filt.words()
            filt.set_sorted_workspace_limit(whole_ws // 4)
            rt = []
            for _ in range(3):
                filt.clear(sh)
                ev = _events(2)
                ev[0].record(stream)
                filt.insert_sequence_async(genome, stream=sh)
                ev[1].record(stream)
                ev[1].synchronize()
                rt.append(ev[0].elapsed_time(ev[1]) / 1e3)
            info = filt.sorted_workspace()
            same = bool(np.array_equal(filt.words(), ref_words))
            t_rng = statistics.median(rt)
            log(f"[bench] insert in {info.last_chunks} window ranges (workspace limit "
                f"{whole_ws / 4 / 2**30:.1f} GiB): {t_rng * 1e3:.2f} ms, words equal: {same}")
            return {"workspace_limit_bytes": whole_ws // 4, "single_insert_workspace_bytes": whole_ws,
                    "ranges": info.last_chunks, "insert_ms": t_rng * 1e3,
                    "insert_kmers_per_s": v_ins / t_rng, "words_equal_single_insert": same}

        ranged = guarded("insert in window ranges", measure_ranges)
        filt.set_sorted_workspace_limit(0)

    placement = None
    if args.placement_check:
        # the same insert and query on further filters, each allocated while the earlier
        # ones are kept (their pages land elsewhere in device memory)
        placement = [{"filter": 0, "insert_ms": ti * 1e3, "query_ms": tq * 1e3}]
        extra = []
        for i in range(args.placement_check):
            g = cs.Filter(cfg, int(wl.bits_per_kmer * genome_windows))
            g.set_insert_path(INSERT_PATHS[args.insert_path])
            gi, gq = timed_insert_query(g, genome, reads, out, sh, max(3, args.steps), 1)
            placement.append({"filter": i + 1, "insert_ms": gi * 1e3, "query_ms": gq * 1e3})
            log(f"[bench] placement: filter {i + 1}: insert {gi * 1e3:.2f} ms, query {gq * 1e3:.2f} ms")
            extra.append(g)
        filt.contains_sequence_async(reads, out, stream=sh)  # the headline filter's answers again
        for g in extra:
            g.close()
        del extra
        torch.cuda.empty_cache()

    positives = int(cs.count_positive(out))
    ins_rate = v_ins / ti
    q_rate = v_q / tq
    value = (v_ins + v_q) / (ti + tq)

    # byte model: insert = input + 64 B per shard run; query = input + hit bytes + 32 B per run
    bytes_ins = genome.numel() + 64 * s_ins
    bytes_q = reads.numel() + out.numel() + 32 * s_q
    dom = "query" if tq >= ti else "insert"
    t_dom, b_dom = (tq, bytes_q) if dom == "query" else (ti, bytes_ins)
    achieved = b_dom / t_dom

    fpr = guarded("fpr", fpr_scenario, genome, reads, wl, out, sh) if args.fpr == "on" else None

    # serving runs on the filter as inserted, before the probes overwrite its buffer
    serving = None
    if args.serve == "on":
        serving = guarded("serving", serving_scenario, filt, genome, reads, wl, args.serve_requests,
                          args.serve_qps)

    ceilings = None
    if args.probes == "on":
        def measure_ceilings() -> dict:
            from cusbf_amd import probes

            log("[bench] measuring memory ceilings")
            words = filt.words_device_ptr()
            sectors = filt.num_shards()
            copy_bytes = min(reads.numel(), 1 << 32)
            copy_bw = probes.stream_copy(reads.data_ptr(), out.data_ptr(), copy_bytes)
            gather = probes.random_gather(words, sectors, min(s_q, 2_000_000_000))
            update = probes.random_sector_update(words, sectors, s_ins)
            # the query's memory traffic without its arithmetic: the reads streamed in and
            # the answers out in 1 KiB tiles, with the query's shard runs as random sectors
            tiles = min(reads.numel(), out.numel()) // 1024
            per_tile = max(1, round(s_q / tiles))
            traffic_s = probes.query_traffic(reads.data_ptr(), out.data_ptr(), tiles, per_tile, words, sectors)
            parts_s = {name: probes.query_traffic(reads.data_ptr(), out.data_ptr(), tiles, per_tile, words,
                                                  sectors, parts) * 1e3
                       for name, parts in (("sectors_only_ms", 0), ("sectors_and_input_ms", 1),
                                           ("sectors_and_output_ms", 2))}
            c = {
                "stream_copy_GBps": copy_bw / 1e9,
                "random_32B_gather_per_s": gather,
                "random_32B_gather_GBps": gather * 32 / 1e9,
                "random_32B_update_per_s": update,
                "buffer_bytes": sectors * 32,
                "query_traffic_probe": {"tiles": tiles, "stream_bytes_each_way": tiles * 1024,
                                        "sectors_per_tile": per_tile, "sectors": tiles * per_tile,
                                        "ms": traffic_s * 1e3, "query_ms_over_probe_ms": tq / traffic_s,
                                        **parts_s},
            }
            log(f"[bench] copy {copy_bw / 1e9:.0f} GB/s, random gather {gather / 1e9:.2f} G/s, "
                f"random update {update / 1e9:.2f} G/s; query traffic alone {traffic_s * 1e3:.2f} ms "
                f"(query {tq * 1e3:.2f} ms)")
            return c

        ceilings = guarded("ceilings", measure_ceilings)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded genome and reads generated on the GPU)",
        "config": {"workload": f"{wl.name}-scale plain GPU path",
                   "genome_bp": wl.genome.total_len, "chromosomes": wl.genome.n_chrom,
                   "reads": wl.reads.n_reads, "read_len": wl.reads.read_len,
                   "k": wl.k, "s": wl.s, "m": wl.m, "hash_count": wl.h,
                   "bits_per_kmer_requested": wl.bits_per_kmer,
                   "filter_bits": filt.filter_bits(),
                   "bits_per_kmer_effective": filt.filter_bits() / max(1, v_ins)},
        "insert_path": {1: "atomic", 2: "sorted"}[filt.insert_path().last_path],
        "insert_kmers_per_s": ins_rate,
        "query_kmers_per_s": q_rate,
        "insert_ms": ti * 1e3,
        "query_ms": tq * 1e3,
        "insert_populated_ms": t_pop * 1e3,
        "insert_populated_kmers_per_s": v_ins / t_pop,
        "insert_in_window_ranges": ranged,
        "insert_shard_runs": s_ins,
        "query_shard_runs": s_q,
        "insert_roofline_frac": bytes_ins / ti / HBM_PEAK,
        "query_roofline_frac": bytes_q / tq / HBM_PEAK,
        "insert_random_sector_frac": (s_ins / ti / ceilings["random_32B_update_per_s"]
                                      if ceilings and "error" not in ceilings else None),
        "query_random_sector_frac": (s_q / tq / ceilings["random_32B_gather_per_s"]
                                     if ceilings and "error" not in ceilings else None),
        "query_positive_kmers": positives,
        "query_positive_frac": positives / max(1, v_q),
        "ceilings": ceilings,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": committed_traffic(wl.name, dom),
                     "kernel": dom, "alg_bytes_per_launch": b_dom},
        "serve_p50_ms": None,
        "serve_p99_ms": None,
        "fpr": fpr,
    }
    if nematode is not None:
        result["nematode"] = nematode
    if placement is not None:
        result["placement_check"] = placement
    if serving is not None:
        result["serve_p50_ms"] = serving.get("server_p50_ms")
        result["serve_p99_ms"] = serving.get("server_p99_ms")
        result["serving"] = serving
    if args.dense == "on":
        result["dense"] = guarded("dense", dense_scenario, genome, wl, sh, args.dense_host_bases)
    if args.fastx == "on":
        result["fastx"] = guarded("fastx", fastx_scenario, genome, reads, wl, args.fastx_bases,
                                  args.fastx_reads)
    if args.batched == "on":
        del out
        filt.close()
        torch.cuda.empty_cache()
        def per_alphabet(path, a):
            return path.replace(".json", f"_{a}.json") if path else ""

        result["batched"] = {
            a: guarded(f"batched {a}", batched_scenario, a, args.batched_records, args.batched_reps,
                       per_alphabet(args.tuning_out, a), args.batched_baselines == "on",
                       per_alphabet(args.tuning_in, a))
            for a in ("dna", "protein")}
    if args.cpu_baseline == "on":
        log("[bench] CPU baseline on a bounded sample")
        result["cpu_baseline"] = guarded("cpu baseline", cpu_baseline, genome, reads, wl, args.cpu_threads)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
