"""Autotuner of the batched record form: launch shape per record-length bucket.

For each bucket (mean record length < 256, < 4096, < 65536, >= 65536 symbols)
the tuner generates a sample batch whose record lengths follow a log-normal
centred in the bucket, captures a graph plan for every candidate launch shape
(kernel variant built for 3 or 4 waves per SIMD, persistent blocks per compute
unit, tiles per work-queue claim or static striding) and times plan replays
with HIP events. Inserts are timed into a scratch filter of the same
configuration and size as the target (so the target filter is never written),
once on the atomic path and once on the sorted path (each path has its own
table entry: the insert kernel is bound by memory-side atomics on one and by
its own arithmetic on the other); queries are timed against that scratch
filter after the insert. The fastest shape per (bucket, op) goes into the
target filter's tuning table.

Usage:
    table = tune(filter)                  # all buckets
    save(table, "tuning.json"); apply(filter, load("tuning.json"))
"""
from __future__ import annotations

import json
import statistics
from dataclasses import asdict, dataclass, field

import numpy as np

from . import batched as bt
from .config import ALPHABET_PROTEIN
from .filter import INSERT_ATOMIC, INSERT_AUTO, INSERT_SORTED, Filter
from .workloads import PROTEIN_LETTERS, uniform_letters

# (median, sigma, min, max) of the sample's record lengths per bucket
BUCKET_SAMPLES = {
    0: (150.0, 0.4, 31, 255),
    1: (1000.0, 0.6, 256, 4095),
    2: (16000.0, 0.6, 4096, 65535),
    3: (200000.0, 0.6, 65536, 2_000_000),
}

BLOCKS_PER_CU = (2, 4, 8, 16, 32)
CLAIMS = (0, 1, 4, 16)
OPS = {"insert": bt.INSERT, "insert_sorted": bt.INSERT_SORTED, "contains": bt.CONTAINS}


@dataclass
class BucketResult:
    bucket: int
    op: str
    best: dict
    best_ms: float
    timings: list = field(default_factory=list)  # (waves, blocks, claim, ms)


def candidates(filt: Filter) -> list:
    """Every launch shape of the filter's batched kernels (none for the generic kernels)."""
    waves = bt.batch_variants(filt)
    return [bt.BatchTuning(w, b, c) for w in waves for b in BLOCKS_PER_CU for c in CLAIMS]


def sample_batch(filt: Filter, bucket: int, symbols: int, seed: int = 1234):
    """(device symbols, device offsets, lengths) of a sample batch for the bucket."""
    import torch

    median, sigma, lo, hi = BUCKET_SAMPLES[bucket]
    rng = np.random.default_rng(seed + bucket)
    lens = []
    total = 0
    while total < symbols:
        x = int(np.clip(round(rng.lognormal(np.log(median), sigma)), lo, hi))
        lens.append(x)
        total += x
    lens = np.array(lens, dtype=np.int64)
    off = np.zeros(lens.size + 1, dtype=np.int64)
    off[1:] = np.cumsum(lens)
    code = filt.config.alphabet.code
    letters = PROTEIN_LETTERS if code == ALPHABET_PROTEIN else filt.config.alphabet.valid_bytes
    w = filt.config.symbol_width
    data = uniform_letters(int(off[-1]) * w, letters, seed + 17 * bucket)
    return data, torch.from_numpy(off).cuda(), lens


def _time_plan(plan: bt.BatchPlan, reps: int, before=None) -> float:
    import torch

    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        if before is not None:
            before()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        plan.launch(stream)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    plan.status(stream)
    return statistics.median(times)


def tune_bucket(filt: Filter, bucket: int, sample_symbols: int = 1 << 28, reps: int = 3,
                scratch: Filter | None = None) -> dict:
    """Times every candidate for both ops on a sample of the bucket; returns
    {op_name: BucketResult} (the table is not changed)."""
    import torch

    data, off, lens = sample_batch(filt, bucket, sample_symbols)
    own_scratch = scratch is None
    if own_scratch:
        scratch = Filter(filt.config, filt.filter_bits())
    k = filt.config.k
    n_win = max(1, data.numel() // filt.config.symbol_width - k + 1)
    out = torch.empty(n_win, dtype=torch.uint8, device="cuda")
    results = {}
    ops = ((bt.INSERT, "insert", INSERT_ATOMIC), (bt.INSERT_SORTED, "insert_sorted", INSERT_SORTED),
           (bt.CONTAINS, "contains", None))
    try:
        for op, name, path in ops:
            timings = []
            if path is not None:
                scratch.set_insert_path(path)
            for t in candidates(filt):
                if path is not None:
                    plan = bt.BatchPlan(scratch, bt.INSERT, data, off, tuning=t)
                    ms = _time_plan(plan, reps, before=lambda: scratch.clear())
                else:
                    plan = bt.BatchPlan(scratch, op, data, off, out, tuning=t)
                    ms = _time_plan(plan, reps)
                plan.close()
                timings.append((t.waves_per_simd, t.blocks_per_cu, t.claim, round(ms, 4)))
                if path is not None:
                    scratch.clear()
            scratch.set_insert_path(INSERT_AUTO)
            if op == bt.INSERT_SORTED:  # leave the sample inserted for the query timings
                with bt.BatchPlan(scratch, bt.INSERT, data, off) as plan:
                    plan.launch()
                    plan.status()
            best = min(timings, key=lambda r: r[3])
            results[name] = BucketResult(bucket, name, asdict(bt.BatchTuning(*best[:3])), best[3],
                                         timings)
    finally:
        scratch.set_insert_path(INSERT_AUTO)
        if own_scratch:
            scratch.close()
    return results


def tune(filt: Filter, buckets=(0, 1, 2, 3), sample_symbols: int = 1 << 28, reps: int = 3,
         log=None) -> dict:
    """Tunes every bucket, writes the winners into the filter's table and returns
    {"bucket<i>": {op: {"best": ..., "best_ms": ..., "timings": [...]}}}."""
    table = {}
    if not candidates(filt):
        return table  # generic kernels: no launch shape to tune
    scratch = Filter(filt.config, filt.filter_bits())
    try:
        for b in buckets:
            res = tune_bucket(filt, b, sample_symbols, reps, scratch=scratch)
            table[f"bucket{b}"] = {op: asdict(r) for op, r in res.items()}
            for op_name, r in res.items():
                bt.set_tuning(filt, b, OPS[op_name], bt.BatchTuning(**r.best))
                if log:
                    log(f"[tune] bucket {b} {op_name}: {r.best} {r.best_ms:.3f} ms")
            scratch.clear()
    finally:
        scratch.close()
    return table


def apply(filt: Filter, table: dict) -> None:
    for key, ops in table.items():
        b = int(key.replace("bucket", ""))
        for op_name, r in ops.items():
            bt.set_tuning(filt, b, OPS[op_name], bt.BatchTuning(**r["best"]))


def save(table: dict, path: str) -> None:
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1)


def load(path: str) -> dict:
    with open(path) as fh:
        return json.load(fh)
