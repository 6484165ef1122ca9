"""Builds libgranddecoder.so (HIP, gfx950) in-tree, and the CPU oracle.

    python library_s_grand_amd/build.py            # product + oracle
    python library_s_grand_amd/build.py --product  # product only

The product is compiled with hipcc for gfx950 only.  Packed-FP32 instruction
selection is disabled (-packed-fp32-ops) and FP contraction is off, so the ML
costs are the sequential fp32 sums of the reference CPU search.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
PRODUCT = os.path.join(HERE, "libgranddecoder.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
SOURCES = ["grand_kernels.hip", "grand_mlp.hip", "grand_host.cpp", "grand_server.cpp"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature; that warning is expected
    noise = [ln for ln in (proc.stdout + proc.stderr).splitlines() if ln.strip() and "packed-fp32-ops" not in ln]
    if proc.returncode != 0:
        sys.stderr.write("\n".join(noise) + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return noise


def build_product(verbose=False) -> str:
    tmp = PRODUCT + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    out = _run(cmd)
    if verbose and out:
        print("\n".join(out))
    os.replace(tmp, PRODUCT)
    return PRODUCT


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE_DIR])
    return os.path.join(ORACLE_DIR, "libgrand_oracle.so")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--product", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    both = not (a.product or a.oracle)
    if a.product or both:
        print(build_product(a.verbose))
    if a.oracle or both:
        print(build_oracle())


if __name__ == "__main__":
    main()
