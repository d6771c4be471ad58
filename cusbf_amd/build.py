"""Builds libcusbf_hip.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: python -m cusbf_amd.build  (or cusbf_amd.build.build()).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "libcusbf_hip.so")
SYNTH_LIB_PATH = os.path.join(HERE, "libcusbf_synth.so")

HIP_SOURCES = ["kernels_wide_dna.hip", "kernels_wide_protein.hip", "kernels_strip.hip", "kernels_strip_batched.hip",
               "kernels_generic.hip", "sorted_insert.hip", "capi.hip",
               "serve.hip", "fastx.hip", "persist.hip", "noalloc.hip"]
SYNTH_SOURCES = ["synth.hip", "probe.hip", "baseline.hip"]

# -packed-fp32-ops: the kernels are integer code, and packed FP32 forms stay off.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _compile(src: str, obj: str) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")


def _link(objs, path: str, libs=()) -> None:
    tmp = path + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, *libs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, path)


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every HIP source (always) and link libcusbf_hip.so and libcusbf_synth.so."""
    from concurrent.futures import ThreadPoolExecutor

    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)

    def objs_of(names):
        return [os.path.join(out_dir, n.replace(".hip", ".o")) for n in names]

    work = [(os.path.join(CSRC, n), o)
            for n, o in zip(HIP_SOURCES + SYNTH_SOURCES, objs_of(HIP_SOURCES + SYNTH_SOURCES))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_compile, s, o) for s, o in work]:
            fut.result()
    _link(objs_of(HIP_SOURCES), LIB_PATH, ["-lz"])  # zlib: gzip FASTX input
    _link(objs_of(SYNTH_SOURCES), SYNTH_LIB_PATH)
    build_cpp_tests()
    build_hip_tests()
    if verbose:
        print(f"built {LIB_PATH}, {SYNTH_LIB_PATH}, {CPP_TEST_PATH} and {HIP_TEST_PATH}")
    return LIB_PATH


REPO = os.path.dirname(HERE)
CPP_TEST_SRC = os.path.join(REPO, "tests", "cpp", "test_filter_api.cpp")
CPP_TEST_PATH = os.path.join(REPO, "tests", "cpp", "_build", "test_filter_api")


def build_cpp_tests() -> str:
    """The C++ API tests: host g++ against include/cusbf/*.hpp and libcusbf_hip.so."""
    os.makedirs(os.path.dirname(CPP_TEST_PATH), exist_ok=True)
    rocm = os.path.dirname(os.path.dirname(hipcc()))
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           f"-I{rocm}/include", f"-I{os.path.join(REPO, 'include')}", CPP_TEST_SRC,
           "-o", CPP_TEST_PATH, f"-L{HERE}", "-lcusbf_hip", f"-L{rocm}/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../../cusbf_amd", f"-Wl,-rpath,{rocm}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{res.stderr[-6000:]}")
    return CPP_TEST_PATH


HIP_TEST_SRC = os.path.join(REPO, "tests", "hip", "test_filter_ref.hip")
HIP_TEST_PATH = os.path.join(REPO, "tests", "hip", "_build", "test_filter_ref")


def build_hip_tests() -> str:
    """A caller's own HIP kernels over filter<Config>::ref() (cusbf/filter_ref_device.hpp)."""
    os.makedirs(os.path.dirname(HIP_TEST_PATH), exist_ok=True)
    cmd = [hipcc(), *[f for f in HIPCC_FLAGS if f not in ("-fPIC", "-std=c++17")], "-std=c++20",
           f"-I{os.path.join(REPO, 'include')}", HIP_TEST_SRC, "-o", HIP_TEST_PATH, f"-L{HERE}",
           "-lcusbf_hip", "-Wl,-rpath,$ORIGIN/../../../cusbf_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"HIP test build failed:\n{res.stderr[-6000:]}")
    return HIP_TEST_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
