"""Build the HIP library libpluto_amd.so for gfx950 (in-tree).

    python -m pluto_amd.build            # or pluto_amd.build.build()

Compiles every ``csrc/*.hip`` with hipcc into ``pluto_amd/libpluto_amd.so``.
Flags:
  --offload-arch=gfx950       MI355X (CDNA4) only
  -ffp-contract=off           no FMA contraction: FP64 results are the IEEE
                              values of the reference's expressions
  -target-feature -packed-fp32-ops
                              no packed FP32 instructions
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpluto_amd.so")
ARCH = os.environ.get("PGPU_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "--offload-arch=" + ARCH,
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I" + os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + ["-x", "hip"] + sources() + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + proc.stdout[-4000:] + proc.stderr[-8000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
