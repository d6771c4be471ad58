"""Ceilings the kernels are measured against (probe.hip): stream copy, random 32-byte
gather, random sector update, and the VALU issue cost of single instructions.

Each memory probe runs `reps` times on the current stream and returns the median
rate measured with HIP events.
"""
from __future__ import annotations

import ctypes
import statistics

from . import workloads as _w

_probe = None


def lib():
    global _probe
    if _probe is None:
        L = _w.synth_lib()
        P, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.cusbf_probe_copy.argtypes = [P, P, u64, P]
        L.cusbf_probe_gather.argtypes = [P, u64, u64, u64, P, P]
        L.cusbf_probe_atomic.argtypes = [P, u64, u64, u64, P]
        L.cusbf_probe_gather_form.argtypes = [P, u64, u64, u64, P, ctypes.c_int, P]
        L.cusbf_probe_alloc.argtypes = [u64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.cusbf_probe_valu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.cusbf_probe_traffic.argtypes = [P, P, u64, ctypes.c_uint32, P, u64, u64, ctypes.c_uint32, P, P]
        for f in (L.cusbf_probe_copy, L.cusbf_probe_gather, L.cusbf_probe_atomic,
                  L.cusbf_probe_gather_form, L.cusbf_probe_alloc, L.cusbf_probe_valu,
                  L.cusbf_probe_traffic):
            f.restype = ctypes.c_int
        _probe = L
    return _probe


def _time(launch, reps: int) -> float:
    import torch

    times = []
    for i in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        rc = launch(i)
        b.record()
        if rc != 0:
            raise RuntimeError(f"probe launch failed (hip error {rc})")
        torch.cuda.synchronize()
        if i:  # first run warms up
            times.append(a.elapsed_time(b) / 1e3)
    return statistics.median(times)


def stream_copy(src_ptr: int, dst_ptr: int, nbytes: int, reps: int = 3) -> float:
    """Copy bandwidth in bytes/s counting read + write bytes."""
    nbytes -= nbytes % 16
    t = _time(lambda i: lib().cusbf_probe_copy(src_ptr, dst_ptr, nbytes, 0), reps)
    return 2 * nbytes / t


def random_gather(buf_ptr: int, num_sectors: int, accesses: int, reps: int = 3) -> float:
    """Random 32-byte sector reads per second."""
    import torch

    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    t = _time(lambda i: lib().cusbf_probe_gather(buf_ptr, num_sectors, accesses, 1000 + i,
                                                 sink.data_ptr(), 0), reps)
    return accesses / t


def query_traffic(in_ptr: int, out_ptr: int, tiles: int, per_tile: int, buf_ptr: int, num_sectors: int,
                  parts: int = 3, reps: int = 3) -> float:
    """Seconds (median of `reps`) for the query's memory traffic without its arithmetic:
    `tiles` 1 KiB tiles streamed in (parts bit 0) and out (bit 1), `per_tile` random
    32-byte sectors per tile."""
    import torch

    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    return _time(lambda i: lib().cusbf_probe_traffic(in_ptr, out_ptr, tiles, per_tile, buf_ptr,
                                                     num_sectors, 3000 + i, parts, sink.data_ptr(), 0),
                 reps)


def random_sector_update(buf_ptr: int, num_sectors: int, accesses: int, reps: int = 3) -> float:
    """Random 32-byte sector updates (four 64-bit atomic ORs) per second."""
    t = _time(lambda i: lib().cusbf_probe_atomic(buf_ptr, num_sectors, accesses, 2000 + i, 0), reps)
    return accesses / t


LOAD_FORMS = ("plain", "nontemporal", "atomic_agent", "atomic_system")
MEMORY_KINDS = ("hipMalloc", "fine-grained", "uncached", "contiguous")


def alloc(nbytes: int, kind: int) -> int:
    """Device buffer of a memory kind (MEMORY_KINDS index); kept until the process ends."""
    p = ctypes.c_void_p()
    rc = lib().cusbf_probe_alloc(nbytes, kind, ctypes.byref(p))
    if rc != 0:
        raise RuntimeError(f"probe allocation failed (hip error {rc})")
    return p.value


def gather_form(buf_ptr: int, num_sectors: int, accesses: int, mode: int, reps: int = 3) -> list:
    """Random 32-byte sector reads per second in load form `mode`, one rate per repetition."""
    import torch

    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    rates = []
    for i in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        rc = lib().cusbf_probe_gather_form(buf_ptr, num_sectors, accesses, 300 + 17 * i + mode,
                                           sink.data_ptr(), mode, 0)
        b.record()
        if rc != 0:
            raise RuntimeError(f"probe launch failed (hip error {rc})")
        torch.cuda.synchronize()
        if i:
            rates.append(accesses / (a.elapsed_time(b) / 1e3))
    return rates


# Instruction kinds of cusbf_probe_valu, in its numbering.
VALU_OPS = ["v_add_u32", "v_or_b32", "v_and_b32", "v_xor_b32", "v_lshrrev_b32", "v_lshlrev_b32",
            "v_bitop3_b32", "v_or3_b32", "v_add3_u32", "v_min_u32", "v_bfe_u32", "v_alignbit_b32",
            "v_perm_b32", "v_lshl_or_b32", "v_mul_lo_u32", "v_mul_hi_u32", "v_mad_u64_u32",
            "v_lshlrev_b64", "v_lshl_add_u64", "v_cmp_lt_u64", "v_mov_b32_dpp", "v_bcnt_u32_b32",
            "v_mov_b32"]


def valu_cycles(op: str, waves_per_simd: int = 8) -> float:
    """Cycles one wave64 instruction of kind `op` holds its SIMD (eight independent
    chains per lane, `waves_per_simd` waves on every SIMD)."""
    out = ctypes.c_double(0.0)
    rc = lib().cusbf_probe_valu(VALU_OPS.index(op), waves_per_simd, ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"valu probe failed (hip error {rc})")
    return out.value


def main(argv=None) -> int:
    """python -m cusbf_amd.probes forms <memory kind> | placement | valu
    forms: the gather in every load form on a 4 GiB buffer of one memory kind;
    placement: plain load form on seven 8 GiB buffers made one after another;
    valu: issue cycles of each instruction kind at 8 and 3 waves per SIMD."""
    import argparse
    import json

    import torch  # noqa: F401  (one HIP runtime per process)

    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["forms", "placement", "valu"])
    ap.add_argument("kind", nargs="?", type=int, default=0)
    ap.add_argument("--accesses", type=int, default=1_000_000_000)
    a = ap.parse_args(argv)
    if a.what == "valu":
        for op in VALU_OPS:
            c8 = [round(valu_cycles(op, 8), 2) for _ in range(3)]
            c3 = [round(valu_cycles(op, 3), 2) for _ in range(3)]
            print(json.dumps({"instruction": op, "cycles_8_waves": c8, "cycles_3_waves": c3}), flush=True)
    elif a.what == "forms":
        sectors = 1 << 27
        buf = alloc(sectors * 32, a.kind)
        for mode, name in enumerate(LOAD_FORMS):
            r = gather_form(buf, sectors, a.accesses, mode)
            print(json.dumps({"memory": MEMORY_KINDS[a.kind], "load": name,
                              "G_sectors_per_s": [round(x / 1e9, 2) for x in r]}), flush=True)
    else:
        G = 1 << 30
        plan = [("hipMalloc, first", 0, 0), ("hipMalloc, second", 0, 0),
                ("9 GiB, 1 GiB-aligned window", 0, 1), ("contiguous", 3, 0),
                ("hipMalloc, third", 0, 0), ("10 GiB, 2 GiB-aligned window", 0, 2),
                ("hipMalloc, fourth", 0, 0)]
        for name, kind, align in plan:
            base = alloc(8 * G + align * G, kind)
            ptr = base if align == 0 else (base + align * G - 1) & ~(align * G - 1)
            r = gather_form(ptr, (8 * G) // 32, a.accesses, 0)
            print(json.dumps({"allocation": name, "G_sectors_per_s": [round(x / 1e9, 2) for x in r]}),
                  flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
