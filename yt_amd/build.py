"""Builds libytrc.so (HIP, gfx950) in-tree with hipcc.

The library is compiled with -ffp-contract=off so that the FP64 ray set-up
and DDA keep the reference's separately rounded IEEE operations (yt's
Cython is compiled by gcc for baseline x86-64, which never fuses), and with
packed FP32 instructions disabled.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libytrc.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("YTRC_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fPIC",
    "-shared",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return [os.path.join(CSRC, "ytrc.hip")]


def build(verbose=False):
    """Compiles every HIP source into yt_amd/libytrc.so (always rebuilds)."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    # The host pass warns that it does not know the packed-fp32 feature.
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
