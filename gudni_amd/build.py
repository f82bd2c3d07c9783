"""Build libgudni_hip.so in-tree with hipcc for gfx950.

Flags that matter for results:
  -ffp-contract=off                          no FMA contraction, host or device
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division on the device
  -fno-gpu-flush-denormals-to-zero           IEEE denormals on the device
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                             no packed FP32 instructions
"""
import glob
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgudni_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "host.cpp"), os.path.join(HERE, "csrc", "raster.hip"),
           os.path.join(HERE, "csrc", "bin.hip"), os.path.join(HERE, "csrc", "adjoint.hip"),
           os.path.join(HERE, "csrc", "serialize.hip")]
ARCH = os.environ.get("GUDNI_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(cand) and os.path.exists(cand):
            return cand
    return "hipcc"


STAMP = LIB + ".stamp"


def _command(out: str):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
            "-o", out] + SOURCES


def fingerprint() -> str:
    """Hash of everything the library is built from: the compile command
    (compiler, target, flags) and the bytes of every source and header."""
    # the command with the tree's own location left out, so that a copy of
    # the tree elsewhere (built here, run there) keeps its fingerprint
    cmd = " ".join(os.path.relpath(a, ROOT) if os.path.isabs(a) and a.startswith(ROOT + os.sep) else a
                   for a in _command("lib"))
    h = hashlib.sha256(cmd.encode())
    files = sorted(set(SOURCES) | set(glob.glob(os.path.join(HERE, "csrc", "*")))
                   | set(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def stale() -> bool:
    """True when the built library is missing or was built from other
    sources or flags than the tree holds now."""
    if not (os.path.exists(LIB) and os.path.exists(STAMP)):
        return True
    with open(STAMP) as fh:
        return fh.read().strip() != fingerprint()


def build(verbose: bool = False) -> str:
    cmd = _command(LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    fp = fingerprint()
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    with open(STAMP + ".tmp", "w") as fh:
        fh.write(fp + "\n")
    os.replace(STAMP + ".tmp", STAMP)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
