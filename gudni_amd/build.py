"""Build libgudni_hip.so in-tree with hipcc for gfx950.

Flags that matter for results:
  -ffp-contract=off                          no FMA contraction, host or device
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division on the device
  -fno-gpu-flush-denormals-to-zero           IEEE denormals on the device
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                             no packed FP32 instructions
"""
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


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
           "-fno-gpu-flush-denormals-to-zero",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           "-o", LIB + ".tmp"] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
