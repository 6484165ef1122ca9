"""Register budget of the search kernel (compiled for gfx950 here, no GPU).

The persistent search kernel runs at one or two waves per SIMD; a spill to
scratch memory (stack traffic per shot) has cost it 20-40 % before, and it
comes from the compiler outlining a per-shot function that takes a struct by
reference or from register pressure.  The only stack it may use is the
16 bytes the out-of-line candidate probes save.
"""
import os
import re
import subprocess

import pytest

import grandtest as T

SRC = os.path.join(T.ROOT, "library_s_grand_amd", "csrc", "grand_kernels.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_search_kernel_does_not_spill(tmp_path):
    asm = tmp_path / "k.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o",
                    str(asm)], check=True, capture_output=True)
    text = asm.read_text()
    kernels = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", text, flags=re.S):
        priv = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", m.group(2)).group(1))
        vgpr = int(re.search(r"\.amdhsa_next_free_vgpr (\d+)", m.group(2)).group(1))
        kernels[m.group(1)] = (priv, vgpr)
    fused = [v for k, v in kernels.items() if "fused_kernel" in k]
    triage = [v for k, v in kernels.items() if "triage_kernel" in k]
    assert fused and triage
    assert fused[0][0] <= 16 and fused[0][1] <= 256, fused
    assert triage[0][0] == 0, triage
