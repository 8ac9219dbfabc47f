"""Builds libytrc.so (HIP, gfx950) in-tree with hipcc.

The library is compiled with -ffp-contract=off so that the FP64 ray set-up
and DDA keep the reference's separately rounded IEEE operations (yt's
Cython is compiled by gcc for baseline x86-64, which never fuses), and with
packed FP32 instructions disabled.

Each translation unit (the host code, and one per family of kernel
variants) compiles to an object in parallel; the objects link into the
shared library.
"""

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

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
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}: {' '.join(cmd[-3:])}")


def build(verbose=False, jobs=None):
    """Compiles every HIP source into yt_amd/libytrc.so (always rebuilds)."""
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    inc = ["-I", os.path.join(ROOT, "include")]
    with tempfile.TemporaryDirectory(prefix="ytrc_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.basename(s)[:-4] + ".o") for s in srcs]
        cmds = [[HIPCC, *FLAGS, *inc, "-c", s, "-o", o] for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    # The host pass warns that it does not know the packed-fp32 feature.
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
