"""Compile the HIP library ``libhz.so`` in-tree for gfx950.

The library is built from ``csrc/hz_capi.hip`` with hipcc; the .so sits next
to this file so that it travels with the source tree.
"""
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = [HERE/"csrc"/"hz_capi.hip"]
HEADERS = sorted((HERE/"csrc").glob("*.hpp")) + [HERE.parent/"include"/"hz.h"]
LIBRARY = HERE/"libhz.so"
ARCH = os.environ.get("HZ_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of the generated code on purpose.
# Multiply-adds are contracted into FMAs only within one source expression
# (-ffp-contract=on; the hot loops spell theirs out), so that the rounding of
# every kernel is set by its source, not by how the compiler schedules it:
# a tile shape, a fused multi-scene launch and a single-scene launch of the
# same row round it identically.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def build(verbose=False):
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *map(str, SOURCES)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIBRARY)
    if verbose:
        print(" ".join(cmd))
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
