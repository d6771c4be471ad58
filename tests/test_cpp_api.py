"""The C++ mirror of cusbf::filter<Config> (include/cusbf/*.hpp), tested by a C++ program.

tests/cpp/test_filter_api.cpp restates the reference's filter tests
(tests/test_bloom_filter.cu, tests/test_query_layout.cpp,
tests/test_normalized_record_batch.cpp) against the C++ API; it links
libcusbf_hip.so and calls the HIP kernels through the C ABI.
"""
from __future__ import annotations

import os
import subprocess

import pytest

from cusbf_amd import build as B


def _binary() -> str:
    if not os.path.exists(B.CPP_TEST_PATH):
        B.build_cpp_tests()
    return B.CPP_TEST_PATH


def test_cpp_host_only_tests():
    res = subprocess.run([_binary(), "--host-only"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failures" in res.stdout


@pytest.mark.gpu
def test_cpp_filter_api_on_gpu(gpu):
    res = subprocess.run([_binary()], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failures" in res.stdout


@pytest.mark.gpu
def test_user_kernels_on_filter_ref(gpu, tmp_path):
    """A caller's own HIP kernels (tests/hip/test_filter_ref.hip) insert and query
    through filter<Config>::ref(): shard_index, sectorizedHashToMasks,
    apply_word_masks and sectorized_contains_packed_kmer (filter_ref.cuh:20-113,
    filter_impl.cuh:18-77). Words and answers equal the oracle's."""
    import numpy as np

    from tests.helpers import make_oracle, random_dna

    if not os.path.exists(B.HIP_TEST_PATH):
        pytest.fail(f"{B.HIP_TEST_PATH} is not built (python -m cusbf_amd.build)")
    ins = random_dna(400_000, seed=71, n_rate=0.001, lower_rate=0.01)
    qry = random_dna(100_000, seed=72, n_rate=0.001) + ins[:100_000]
    (tmp_path / "ins").write_bytes(ins)
    (tmp_path / "qry").write_bytes(qry)
    prefix = str(tmp_path / "out")
    res = subprocess.run([B.HIP_TEST_PATH, str(tmp_path / "ins"), str(tmp_path / "qry"), prefix],
                         capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr
    words = np.fromfile(prefix + ".words", dtype=np.uint64)
    hits = np.fromfile(prefix + ".hits", dtype=np.uint8)
    o = make_oracle(31, 28, 16, 4)
    w = o.new_filter(words.size // 4)
    o.insert(ins, w)
    assert np.array_equal(words, w)
    assert np.array_equal(hits, o.contains(qry, w))
    assert 0 < hits.sum() < hits.size
