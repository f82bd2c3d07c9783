"""Build recipe for the oracle's C restatement (test infrastructure only).

Output: oracle/liboracle_gudni.so, compiled with gcc -O2 -ffp-contract=off so
every FP32 operation happens as written (IEEE binary32 via SSE, no FMA).

The reference's own kernels (src/Graphics/Gudni/OpenCL/Kernels.cl) are not
built here: they need the #define block the Haskell host generates
(OpenCL/Setup.hs:45-64, CppDefines.hs:66-70), an edit of their first source
line (a stray token the host's line replacement removes) and an OpenCL
builtin runtime for the CPU, none of which exist in this image.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_gudni.so")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(HERE, "gudni_oracle.c"), os.path.join(HERE, "gudni_diff.c")]
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-ffp-contract=off",
           "-fno-fast-math", "-fPIC", "-shared", "-o", LIB + ".tmp"] + srcs + ["-lm", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
