"""Memory-system ceilings (probe.hip): stream copy, random 32-byte gather, random sector update.

Each function runs its probe `reps` times on the current stream and returns the
median rate measured with HIP events.
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
        for f in (L.cusbf_probe_copy, L.cusbf_probe_gather, L.cusbf_probe_atomic):
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


def random_sector_update(buf_ptr: int, num_sectors: int, accesses: int, reps: int = 3) -> float:
    """Random 32-byte sector updates (four 64-bit atomic ORs) per second."""
    t = _time(lambda i: lib().cusbf_probe_atomic(buf_ptr, num_sectors, accesses, 2000 + i, 0), reps)
    return accesses / t
