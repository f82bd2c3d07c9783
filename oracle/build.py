"""Build recipe for the oracle's C restatement (test infrastructure only).

Output: oracle/liboracle_gudni.so, compiled with gcc -O2 -ffp-contract=off so
every FP32 operation happens as written (IEEE binary32 via SSE, no FMA).

The reference's own kernels (src/Graphics/Gudni/OpenCL/Kernels.cl) are not
built here: they need the #define block the Haskell host generates
(OpenCL/Setup.hs:45-64, CppDefines.hs:66-70), an edit of their first source
line (a stray token the host's line replacement removes) and an OpenCL
builtin runtime for the CPU, none of which exist in this image.
"""
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_gudni.so")


STAMP = LIB + ".stamp"
SOURCES = [os.path.join(HERE, "gudni_oracle.c"), os.path.join(HERE, "gudni_diff.c"),
           os.path.join(HERE, "gudni_oracle.h")]


def _command(out: str):
    return (["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-ffp-contract=off",
             "-fno-fast-math", "-fPIC", "-shared", "-o", out] + SOURCES[:2] + ["-lm", "-lpthread"])


def fingerprint() -> str:
    """Hash of the compile command and the bytes of every source."""
    # the command with the tree's own location left out, so that a copy of
    # the tree elsewhere keeps its fingerprint
    cmd = " ".join(os.path.relpath(a, HERE) if os.path.isabs(a) and a.startswith(HERE + os.sep) else a
                   for a in _command("lib"))
    h = hashlib.sha256(cmd.encode())
    for f in SOURCES:
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode())
            h.update(fh.read())
    return h.hexdigest()


def stale() -> bool:
    """True when the library is missing or built from other sources/flags."""
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
