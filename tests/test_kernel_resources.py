"""The search kernel as compiled for gfx950 (here, no GPU).

* Register budget: the persistent search kernel runs at one or two waves per
  SIMD; a spill to scratch memory (stack traffic per shot) has cost it 20-40 %
  before, and it comes from the compiler outlining a per-shot function that
  takes a struct by reference or from register pressure.  The only stack it may
  use is the 16 bytes the out-of-line candidate probes save.
* Publication order of the shared-task protocol: the task's fields are written
  with device-scope stores and published by a relaxed exchange on its ctl word;
  a chunk's results are device-scope atomics counted done by a relaxed add; the
  slot is freed by a relaxed store.  Each of these must come after a wait for
  every earlier memory operation of the wave (s_waitcnt vmcnt(0): gfx9 counts
  stores in vmcnt), with no other memory operation and no branch target in
  between.
"""
import os
import re
import subprocess

import pytest

import grandtest as T

SRC = os.path.join(T.ROOT, "library_s_grand_amd", "csrc", "grand_kernels.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def kernel_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    asm = tmp_path_factory.mktemp("asm") / "k.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o",
                    str(asm)], check=True, capture_output=True)
    return asm.read_text()


def test_search_kernel_does_not_spill(kernel_asm):
    text = kernel_asm
    kernels = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", text, flags=re.S):
        priv = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", m.group(2)).group(1))
        vgpr = int(re.search(r"\.amdhsa_next_free_vgpr (\d+)", m.group(2)).group(1))
        kernels[m.group(1)] = (priv, vgpr)
    # the search kernel is compiled twice: first-hit and ML
    fused = [v for k, v in kernels.items() if "fused_kernel" in k]
    triage = [v for k, v in kernels.items() if "triage_kernel" in k]
    assert len(fused) == 2 and triage
    assert all(priv <= 16 and vgpr <= 256 for priv, vgpr in fused), fused
    assert triage[0][0] == 0, triage


VMEM = re.compile(r"^\s*(flat|global|buffer|scratch)_(load|store|atomic)")
LABEL = re.compile(r"^[.\w$]+:")


def _waits_before(lines, i):
    """True when, scanning back from line i, an s_waitcnt with vmcnt(0) comes
    before any vector memory operation, label or function start."""
    for j in range(i - 1, -1, -1):
        ln = lines[j].split(";")[0].rstrip()
        if not ln.strip():
            continue
        if ln.strip().startswith("s_waitcnt") and "vmcnt(0)" in ln:
            return True
        if VMEM.match(ln) or LABEL.match(ln):
            return False
    return False


def test_task_publication_waits_for_memory(kernel_asm):
    lines = kernel_asm.splitlines()
    # TaskSlot: ctl at byte 0 (publishing exchange, no returned value), done at
    # 24 (a chunk counted done), busy at 28 (the slot freed, device scope)
    publish = [i for i, ln in enumerate(lines) if re.match(r"^\s*(flat|global)_atomic_swap_x2 v\[\d+:\d+\], v\[\d+:\d+\]\s*$", ln)]
    done = [i for i, ln in enumerate(lines) if re.match(r"^\s*(flat|global)_atomic_add v\[\d+:\d+\], v\d+ offset:24\s*$", ln)]
    free = [i for i, ln in enumerate(lines) if re.match(r"^\s*(flat|global)_store_dword v\[\d+:\d+\], v\d+ offset:28 sc1\s*$", ln)]
    assert len(publish) >= 3 and len(done) >= 3 and len(free) >= 2, (len(publish), len(done), len(free))
    bad = [lines[i].strip() + f" (line {i + 1})" for i in publish + done + free if not _waits_before(lines, i)]
    assert not bad, bad
