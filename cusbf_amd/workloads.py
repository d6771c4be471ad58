"""Seeded synthetic workloads of the benchmark configurations, generated on the GPU.

The host plans chromosome lengths and repeat copies with numpy (small arrays);
libcusbf_synth.so writes every base on the device as a pure function of
(seed, position). Chromosomes form one record batch joined by single 'N'
separators; reads are fixed-length records joined by 'N'.

Configurations (sizes as in the benchmark definitions):
  small     10 Mbp uniform ACGT; query it and an independent 10 Mbp sequence
  nematode  100 Mbp, 35% GC, 6 chromosomes of 12-21 Mbp, 5% tandem and
            interspersed repeats at 5% divergence, 85.7 bits/k-mer;
            20M 150 bp reads, half from the genome with 0.5% substitutions
  human     3.1 Gbp, 41% GC, 24 chromosomes of 50-250 Mbp, 45% repeat copies of
            300 bp-6 kbp at 2-15% divergence, 21.4 bits/k-mer; 100M 150 bp
            reads, half from the genome with 1% substitutions
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import build as _build

REPEAT_DTYPE = np.dtype([("dst", "<u8"), ("len", "<u4"), ("family", "<u4"), ("offset", "<u4"),
                         ("period", "<u4"), ("div_ppm", "<u4"), ("pad", "<u4")])

_synth = None


def synth_lib():
    global _synth
    if _synth is None:
        path = _build.SYNTH_LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build with python -m cusbf_amd.build")
        from ._lib import share_hip_runtime

        share_hip_runtime()
        L = ctypes.CDLL(path)
        P, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.cusbf_synth_genome.argtypes = [u64, u32, P, u32, P, u64, P, u64, P]
        L.cusbf_synth_genome.restype = ctypes.c_int
        L.cusbf_synth_reads.argtypes = [u64, P, u64, u64, u64, u32, u32, u32, P, P]
        L.cusbf_synth_reads.restype = ctypes.c_int
        L.cusbf_synth_repeat_copy_size.restype = u64
        L.cusbf_synth_long_reads.argtypes = [u64, P, u64, u64, P, u32, P, P]
        L.cusbf_synth_long_reads.restype = ctypes.c_int
        L.cusbf_synth_letters.argtypes = [u64, ctypes.c_char_p, u32, P, u64, P]
        L.cusbf_synth_letters.restype = ctypes.c_int
        L.cusbf_bench_each_record.argtypes = [P, ctypes.c_int, P, P, u32, P, u64, u32, P, P]
        L.cusbf_bench_each_record.restype = ctypes.c_int
        assert L.cusbf_synth_repeat_copy_size() == REPEAT_DTYPE.itemsize
        _synth = L
    return _synth


@dataclass
class GenomeSpec:
    name: str
    total_len: int
    n_chrom: int
    chrom_min: int
    chrom_max: int
    gc: float
    repeat_frac: float
    repeat_len: tuple = (300, 6000)
    divergence: tuple = (0.02, 0.15)
    tandem_frac: float = 0.0     # fraction of repeat bases in tandem arrays
    n_families: int = 500
    seed: int = 1


@dataclass
class ReadSpec:
    n_reads: int
    read_len: int = 150
    genome_frac: float = 0.5
    substitution: float = 0.01
    seed: int = 2


@dataclass
class Workload:
    name: str
    k: int = 31
    s: int = 28
    m: int = 16
    h: int = 4
    bits_per_kmer: float = 21.4
    genome: Optional[GenomeSpec] = None
    reads: Optional[ReadSpec] = None
    extra: dict = field(default_factory=dict)


WORKLOADS = {
    "nematode": Workload(
        "nematode", bits_per_kmer=85.7,
        genome=GenomeSpec("nematode", 100_000_000, 6, 12_000_000, 21_000_000, 0.35, 0.05,
                          repeat_len=(300, 6000), divergence=(0.05, 0.05), tandem_frac=0.5,
                          seed=11),
        reads=ReadSpec(20_000_000, 150, 0.5, 0.005, seed=12)),
    "human": Workload(
        "human", bits_per_kmer=21.4,
        genome=GenomeSpec("human", 3_100_000_000, 24, 50_000_000, 250_000_000, 0.41, 0.45,
                          repeat_len=(300, 6000), divergence=(0.02, 0.15), seed=21),
        reads=ReadSpec(100_000_000, 150, 0.5, 0.01, seed=22)),
}


def chromosome_lengths(spec: GenomeSpec) -> np.ndarray:
    """n_chrom lengths in [chrom_min, chrom_max] summing to total_len."""
    rng = np.random.default_rng(spec.seed)
    lens = rng.uniform(spec.chrom_min, spec.chrom_max, spec.n_chrom)
    for _ in range(64):  # rescale the free part, clip to the bounds, repeat
        residual = spec.total_len - lens.sum()
        if abs(residual) < 1:
            break
        room = (spec.chrom_max - lens) if residual > 0 else (lens - spec.chrom_min)
        if room.sum() <= 0:
            break
        lens = np.clip(lens + residual * room / room.sum(), spec.chrom_min, spec.chrom_max)
    lens = np.floor(lens).astype(np.int64)
    lens[np.argmax(lens)] += spec.total_len - int(lens.sum())
    return lens


def plan_repeats(spec: GenomeSpec) -> np.ndarray:
    """Disjoint repeat copies covering repeat_frac of the genome, sorted by position."""
    rng = np.random.default_rng(spec.seed + 1000)
    target = int(spec.repeat_frac * spec.total_len)
    if target == 0:
        return np.zeros(0, dtype=REPEAT_DTYPE)
    lo, hi = spec.repeat_len
    fam_len = np.exp(rng.uniform(np.log(lo), np.log(hi), spec.n_families)).astype(np.int64)
    tandem_target = int(spec.tandem_frac * target)
    # interspersed copies: fragments of family consensus sequences
    est = max(16, int(1.3 * (target - tandem_target) / max(1.0, 0.75 * fam_len.mean())) + 16)
    fam = rng.integers(0, spec.n_families, est)
    length = np.maximum(lo // 2, (fam_len[fam] * rng.uniform(0.5, 1.0, est)).astype(np.int64))
    keep = np.cumsum(length) <= target - tandem_target
    fam, length = fam[keep], length[keep]
    offset = ((fam_len[fam] - length).clip(min=0) * rng.random(fam.size)).astype(np.int64)
    period = np.zeros(fam.size, dtype=np.int64)
    # tandem arrays: a 2-100 bp unit repeated over 100-3000 bp
    if tandem_target:
        n_t = max(1, int(tandem_target / 1550) + 16)
        t_len = rng.integers(100, 3000, n_t)
        keep = np.cumsum(t_len) <= tandem_target
        t_len = t_len[keep]
        fam = np.concatenate([fam, spec.n_families + rng.integers(0, 1 << 20, t_len.size)])
        length = np.concatenate([length, t_len])
        offset = np.concatenate([offset, np.zeros(t_len.size, dtype=np.int64)])
        period = np.concatenate([period, rng.integers(2, 101, t_len.size)])
    n = length.size
    order = rng.permutation(n)
    fam, length, offset, period = fam[order], length[order], offset[order], period[order]
    div = rng.uniform(spec.divergence[0], spec.divergence[1], n)
    # place copies in order with random gaps summing to the unique sequence length
    free = spec.total_len - int(length.sum())
    cuts = np.sort(rng.integers(0, free + 1, n))
    gaps = np.diff(np.concatenate([[0], cuts]))
    dst = np.cumsum(gaps) + np.concatenate([[0], np.cumsum(length)[:-1]])
    out = np.zeros(n, dtype=REPEAT_DTYPE)
    out["dst"] = dst
    out["len"] = length
    out["family"] = fam
    out["offset"] = offset
    out["period"] = period
    out["div_ppm"] = np.round(div * 1e6).astype(np.uint32)
    return out


def generate_genome(spec: GenomeSpec, device="cuda"):
    """(device uint8 buffer with 'N' between chromosomes, chromosome lengths)."""
    import torch

    lens = chromosome_lengths(spec)
    starts = np.concatenate([[0], np.cumsum(lens + 1)[:-1]]).astype(np.uint64)
    buffer_len = int(lens.sum()) + spec.n_chrom - 1
    reps = plan_repeats(spec)
    d_starts = torch.from_numpy(starts.view(np.int64)).to(device)
    d_reps = torch.from_numpy(np.frombuffer(reps.tobytes() or b"\0" * 32, dtype=np.uint8).copy()).to(device)
    out = torch.empty(buffer_len, dtype=torch.uint8, device=device)
    rc = synth_lib().cusbf_synth_genome(spec.seed, int(spec.gc * 1e6), d_starts.data_ptr(),
                                        spec.n_chrom, d_reps.data_ptr(), reps.size,
                                        out.data_ptr(), buffer_len, 0)
    if rc != 0:
        raise RuntimeError(f"genome generation launch failed (hip error {rc})")
    torch.cuda.synchronize()
    return out, lens


def generate_reads(spec: ReadSpec, genome, device="cuda", gc: float = 0.5):
    """Device buffer of spec.n_reads records of read_len bases joined by 'N'."""
    import torch

    total = spec.n_reads * (spec.read_len + 1) - 1
    out = torch.empty(total, dtype=torch.uint8, device=device)
    genome_reads = int(spec.n_reads * spec.genome_frac)
    rc = synth_lib().cusbf_synth_reads(spec.seed, genome.data_ptr(), genome.numel(),
                                       spec.n_reads, genome_reads, spec.read_len,
                                       int(spec.substitution * 1e6), int(gc * 1e6),
                                       out.data_ptr(), 0)
    if rc != 0:
        raise RuntimeError(f"read generation launch failed (hip error {rc})")
    torch.cuda.synchronize()
    return out


def uniform_dna(n: int, seed: int, device="cuda"):
    """n uniform random ACGT bases on the device."""
    spec = GenomeSpec("uniform", n, 1, n, n, 0.5, 0.0, seed=seed)
    return generate_genome(spec, device)[0]


def generate_long_reads(genome, num_reads: int, min_len: int = 10_000, max_len: int = 50_000,
                        error: float = 0.05, seed: int = 5, device="cuda"):
    """(device buffer of long reads joined by 'N', host offsets of the reads).

    Lengths are uniform in [min_len, max_len]; each base carries a mixed
    substitution / insertion / deletion error with total probability `error`.
    """
    import torch

    rng = np.random.default_rng(seed)
    lens = rng.integers(min_len, max_len + 1, num_reads).astype(np.int64)
    offsets = np.zeros(num_reads + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(lens + 1)
    offsets[-1] -= 1  # no separator after the last read
    if genome.numel() <= 2 * max_len:
        raise ValueError("genome too short for the requested read lengths")
    d_off = torch.from_numpy(offsets).to(device)
    out = torch.empty(int(offsets[-1]), dtype=torch.uint8, device=device)
    rc = synth_lib().cusbf_synth_long_reads(seed, genome.data_ptr(), genome.numel(), num_reads,
                                            d_off.data_ptr(), int(error * 1e6), out.data_ptr(), 0)
    if rc != 0:
        raise RuntimeError(f"long-read generation launch failed (hip error {rc})")
    torch.cuda.synchronize()
    return out, offsets


# --- batched uneven-length records -------------------------------------------------

DNA_LETTERS = b"ACGT"
PROTEIN_LETTERS = b"ACDEFGHIKLMNPQRSTVWY"  # the 20 standard residues


@dataclass
class BatchedSpec:
    """Many records of log-normal length, back to back (the batched record form)."""

    name: str = "lognormal-1M"
    records: int = 1_000_000
    median: float = 1000.0
    sigma: float = 1.5
    min_len: int = 31
    max_len: int = 2_000_000
    letters: bytes = DNA_LETTERS
    seed: int = 31


def batched_lengths(spec: BatchedSpec) -> np.ndarray:
    rng = np.random.default_rng(spec.seed)
    lens = np.round(rng.lognormal(np.log(spec.median), spec.sigma, spec.records))
    return np.clip(lens, spec.min_len, spec.max_len).astype(np.int64)


def uniform_letters(n: int, letters: bytes, seed: int, device="cuda"):
    """n bytes drawn uniformly from `letters`, generated on the device."""
    import torch

    out = torch.empty(max(1, n), dtype=torch.uint8, device=device)
    rc = synth_lib().cusbf_synth_letters(seed, letters, len(letters), out.data_ptr(), n, 0)
    if rc != 0:
        raise RuntimeError(f"letters generator failed: hip error {rc}")
    return out[:n]


def generate_batched(spec: BatchedSpec, device="cuda"):
    """(device symbols, device int64 offsets, host lengths) of the spec's records."""
    import torch

    lens = batched_lengths(spec)
    off = np.zeros(lens.size + 1, dtype=np.int64)
    off[1:] = np.cumsum(lens)
    data = uniform_letters(int(off[-1]), spec.letters, spec.seed, device)
    return data, torch.from_numpy(off).to(device), lens


def record_index(lens: np.ndarray, device="cuda"):
    """Device int64 array: the record of every symbol."""
    import torch

    d_lens = torch.from_numpy(lens).to(device)
    ids = torch.arange(lens.size, device=device, dtype=torch.int64)
    return torch.repeat_interleave(ids, d_lens, output_size=int(lens.sum()))


def mutate_odd_records(data, lens: np.ndarray, letters: bytes, seed: int):
    """Same lengths; every odd record replaced by fresh random letters."""
    import torch

    fresh = uniform_letters(data.numel(), letters, seed, data.device)
    parity = torch.from_numpy((np.arange(lens.size) & 1).astype(np.uint8)).to(data.device)
    odd = torch.repeat_interleave(parity, torch.from_numpy(lens).to(data.device),
                                  output_size=data.numel()).bool()
    return torch.where(odd, fresh, data)


def join_records(data, lens: np.ndarray, separator: int, rec_idx=None):
    """Device copy of the records joined by one separator byte (the normalised
    record batch of a width-1 alphabet)."""
    import torch

    n = data.numel()
    if rec_idx is None:
        rec_idx = record_index(lens, data.device)
    out = torch.full((n + max(0, lens.size - 1),), separator, dtype=torch.uint8, device=data.device)
    step = 1 << 28
    for a in range(0, n, step):
        b = min(n, a + step)
        pos = torch.arange(a, b, device=data.device, dtype=torch.int64) + rec_idx[a:b]
        out[pos] = data[a:b]
    return out


# --- serving requests ------------------------------------------------------------------

def serving_requests(short: np.ndarray, read_len: int, long_pool: np.ndarray,
                     long_off: np.ndarray, n_requests: int, seed: int,
                     max_reads: int = 1_000_000):
    """Requests of the serving configuration, as (host bytes, (n, 2) uint64 ranges).

    Sizes are log-uniform over 1..max_reads reads of read_len bases. Each request
    is, with probability 1/2, a run of consecutive short reads from `short` (reads
    of read_len joined by one separator) or enough consecutive long reads of
    `long_pool` (reads at `long_off`, as generate_long_reads returns them) to hold
    at least the same number of bases (at least one read, at most the pool).
    """
    rng = np.random.default_rng(seed)
    stride = read_len + 1
    pool_reads = (short.size + 1) // stride
    n_long = long_off.size - 1
    long_len = np.diff(long_off) - 1
    long_len[-1] += 1  # no separator after the last read
    units = np.exp(rng.uniform(0, np.log(max_reads), n_requests)).astype(np.int64)
    kinds = rng.random(n_requests) < 0.5
    reqs = []
    for i in range(n_requests):
        if kinds[i]:
            n = int(min(units[i], pool_reads))
            st = int(rng.integers(0, pool_reads - n + 1))
            seq = short[st * stride:(st + n) * stride - 1]
            rg = np.stack([np.arange(n, dtype=np.uint64) * stride,
                           np.full(n, read_len, dtype=np.uint64)], axis=1)
        else:
            target = int(units[i]) * read_len
            st = int(rng.integers(0, max(1, n_long - 64)))
            c = int(np.searchsorted(np.cumsum(long_len[st:]), target)) + 1
            c = max(1, min(c, n_long - st))
            base = int(long_off[st])
            seq = long_pool[base:int(long_off[st + c - 1]) + int(long_len[st + c - 1])]
            rg = np.stack([(long_off[st:st + c] - base).astype(np.uint64),
                           long_len[st:st + c].astype(np.uint64)], axis=1)
        reqs.append((seq, rg))
    return reqs
