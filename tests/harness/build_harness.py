"""Builds the test harnesses (tests/harness/ac_hermite: shared-step Hermite
over the legacy ABI; tests/harness/ac_full: two-time-scale Ahmad-Cohen over
the device-resident entry points) against the HIP library and the oracle
library, both in-tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ac_hermite")
NAMES = ("ac_hermite", "ac_full")


def build(verbose=False):
    for name in NAMES:
        out = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, name + ".cpp"),
               "-L", os.path.join(ROOT, "abyss_amd"), "-labyss_regf",
               "-L", os.path.join(ROOT, "oracle"), "-loracle",
               "-Wl,-rpath,$ORIGIN/../../abyss_amd:$ORIGIN/../../oracle", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
