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
