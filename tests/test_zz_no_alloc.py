"""The no-allocation batched query (cusbf_contains_batched_noalloc_async), through the C ABI.

Its answers must equal cusbf_contains_batched_async's byte for byte on the same
batch: two calls in a row on one GPU, the second with a smaller batch (the
shared scratch keeps the first call's bitmap, which the second must clear),
for a lane-strip DNA configuration, the protein front end, a configuration of
the runtime-span lane strips and one of the generic kernels (DNA triplets). Malformed offsets are reported by cusbf_batched_status, and a
batch above the scratch's limit is refused before anything is launched.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.helpers import make_config
from tests.test_gpu_batched import (device, device_offsets, lognormal_lengths, offsets_of,
                                    symbols_for)

pytestmark = pytest.mark.gpu


def _lib():
    from cusbf_amd import _lib

    return _lib.load()


def _query(lib, fn, f, d_data, n_sym, d_off, n_rec, out):
    rc = fn(f.handle, d_data.data_ptr(), n_sym, d_off.data_ptr(), n_rec, out.data_ptr(),
            out.numel(), None)
    return rc


@pytest.mark.parametrize("alphabet,k,s,m,h", [("dna", 31, 28, 16, 4), ("protein", 10, 8, 6, 4),
                                              ("dna", 27, 25, 13, 4), ("triplet", 7, 5, 4, 4)])
def test_noalloc_equals_batched_query_twice(gpu, alphabet, k, s, m, h):
    import torch

    import cusbf_amd as cs
    from cusbf_amd import batched as bt

    lib = _lib()
    cfg = make_config(k, s, m, h, alphabet)
    f = cs.Filter(cfg, 1 << 24)
    big_lens = lognormal_lengths(3000, 800, 1.5, seed=k + 1)
    big_lens[:3] = [0, k - 1, k]
    small_lens = lognormal_lengths(400, 300, 1.2, seed=k + 2)
    inserted = symbols_for(alphabet, int(big_lens.sum()), seed=k + 3)
    big_off = offsets_of(big_lens)
    bt.insert_batched_async(f, device(inserted), device_offsets(big_off))
    bt.batched_status(f)
    # the large batch: half of it the inserted records (positives), half fresh
    for lens, seed in ((big_lens, k + 4), (small_lens, k + 5)):
        off = offsets_of(lens)
        n = int(off[-1])
        data = symbols_for(alphabet, n, seed=seed)
        if lens is big_lens:
            data = inserted[: n // 2] + data[n // 2:]
        d_data, d_off = device(data), device_offsets(off)
        windows = n - k + 1
        want = torch.full((windows,), 7, dtype=torch.uint8, device="cuda")
        got = torch.full((windows,), 9, dtype=torch.uint8, device="cuda")
        assert _query(lib, lib.cusbf_contains_batched_async, f, d_data, n, d_off, len(lens), want) == 0
        assert lib.cusbf_batched_status(f.handle, None) == 0
        assert _query(lib, lib.cusbf_contains_batched_noalloc_async, f, d_data, n, d_off,
                      len(lens), got) == 0
        assert lib.cusbf_batched_status(f.handle, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"{alphabet} k={k}: {int((got != want).sum())} answers differ"
        assert 0 < int(got.sum()) < windows or lens is small_lens


def test_noalloc_malformed_offsets_and_limit(gpu):
    import torch

    import cusbf_amd as cs
    from cusbf_amd.errors import ErrorCategory

    lib = _lib()
    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    data = device(symbols_for("dna", 5000, seed=3))
    bad = device_offsets(np.array([0, 3000, 2000, 5000], np.uint64))
    out = torch.zeros(5000 - 30, dtype=torch.uint8, device="cuda")
    assert _query(lib, lib.cusbf_contains_batched_noalloc_async, f, data, 5000, bad, 3, out) == 0
    st = lib.cusbf_batched_status(f.handle, None)
    assert st == 1 + int(ErrorCategory.invalid_argument)
    limit = (1 << 28) - 4096
    st = lib.cusbf_contains_batched_noalloc_async(f.handle, data.data_ptr(), limit + 1,
                                                  bad.data_ptr(), 3, out.data_ptr(), out.numel(), None)
    assert st == 1 + int(ErrorCategory.invalid_argument)
    buf = ctypes.create_string_buffer(256)
    lib.cusbf_last_error(buf, 256)
    assert b"no-allocation scratch" in buf.value
