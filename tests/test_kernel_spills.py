"""No kernel of the library spills registers to scratch memory.

Work-queue batched query kernels whose spilled values were reloaded while other
vector memory operations were still in flight returned different answers from
run to run on MI355X. Every kernel is therefore built with a register budget
under which it does not spill at all. This test compiles the kernel sources for
gfx950 (no GPU needed) and checks every kernel in two ways: the compiler's
resource-usage report must give a private segment (ScratchSize) of 0 bytes per
lane, and the assembly must hold no scratch access (so no `s_waitcnt vmcnt(N)`
with N > 0 can be issued while a scratch reload is outstanding).
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cusbf_amd", "csrc")
SOURCES = ["kernels_strip.hip", "kernels_strip_batched.hip", "kernels_wide_dna.hip",
           "kernels_wide_protein.hip", "kernels_generic.hip",
           "sorted_insert.hip", "synth.hip", "probe.hip"]


def _hipcc():
    from cusbf_amd import build

    try:
        return build.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")


def kernel_bodies(asm: str):
    # function labels carry the "; @name" comment; data objects (constant tables) do not
    for m in re.finditer(r"\n(_Z\w+):[ \t]*; @[^\n]*\n", asm):
        end = asm.find(".Lfunc_end", m.end())
        yield m.group(1), [line.strip() for line in asm[m.end():end].split("\n")]


def partial_waits_with_reload_pending(body) -> int:
    pending = False
    bad = 0
    for line in body:
        if line.startswith("scratch_load"):
            pending = True
        w = re.match(r"s_waitcnt\s+(.*)", line)
        if w:
            vm = re.search(r"vmcnt\((\d+)\)", w.group(1))
            if vm and int(vm.group(1)) == 0:
                pending = False
            elif vm and pending:
                bad += 1
    return bad


def scratch_sizes(remarks: str) -> dict:
    """Kernel name -> ScratchSize [bytes/lane] from -Rpass-analysis=kernel-resource-usage."""
    sizes, name = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            sizes[name] = int(m.group(1))
    return sizes


@pytest.mark.parametrize("source", SOURCES)
def test_no_kernel_spills(source):
    from cusbf_amd import build

    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "k.s")
        flags = [f for f in build.HIPCC_FLAGS if f not in ("-fPIC", "-Wall")]
        cmd = [_hipcc(), *flags, "--cuda-device-only", "-S", os.path.join(CSRC, source), "-o", out,
               "-Rpass-analysis=kernel-resource-usage"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-3000:]
        asm = open(out).read()
    sizes = scratch_sizes(res.stderr)
    kernels = [name for name, _ in kernel_bodies(asm)]
    assert kernels and set(kernels) <= set(sizes), "resource-usage remarks missing for some kernels"
    spilling = {name: n for name, n in sizes.items() if n != 0}
    assert not spilling, spilling
    scratch = {name for name, body in kernel_bodies(asm)
               if any(line.startswith(("scratch_", "buffer_store_dword off")) for line in body)}
    assert not scratch, scratch
    offenders = {name: n for name, body in kernel_bodies(asm)
                 if (n := partial_waits_with_reload_pending(body))}
    assert not offenders, offenders
