"""Build recipe for the oracle's C restatement (test infrastructure only).

Compiles oracle/regf_oracle.c into oracle/liboracle.so with gcc.  The flags
keep the arithmetic a plain IEEE-754 FP64 restatement (no FMA contraction)
and use OpenMP over targets so the same library doubles as the multi-core
CPU baseline.  x86-64-v3 (AVX2/FMA-capable hosts) keeps the binary portable
across the build container and the GPU host.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "regf_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-fPIC", "-shared", "-o", OUT + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
