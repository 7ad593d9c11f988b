"""Builds the Hermite test harness (tests/harness/ac_hermite) against the
HIP library and the oracle library, both in-tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ac_hermite")


def build(verbose=False):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "ac_hermite.cpp"),
           "-L", os.path.join(ROOT, "abyss_amd"), "-labyss_regf",
           "-L", os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../abyss_amd:$ORIGIN/../../oracle", "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
