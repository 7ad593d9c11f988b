"""Builds the test harnesses against the HIP library in-tree:
tests/harness/ac_hermite (shared-step Hermite over the legacy ABI) and
tests/harness/ac_full (two-time-scale Ahmad-Cohen over the device-resident
entry points), which also link the oracle library as their reference
provider, and tests/harness/legacy_timing (the legacy calling pattern timed
from C, for bench.py), which links the HIP library only."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ac_hermite")
# harness -> links the oracle library
NAMES = {"ac_hermite": True, "ac_full": True, "legacy_timing": False}


def build(verbose=False):
    for name, with_oracle in NAMES.items():
        out = os.path.join(HERE, name)
        oracle = ["-L", os.path.join(ROOT, "oracle"), "-loracle"] if with_oracle else []
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, name + ".cpp"),
               "-L", os.path.join(ROOT, "abyss_amd"), "-labyss_regf", *oracle,
               "-Wl,-rpath,$ORIGIN/../../abyss_amd:$ORIGIN/../../oracle", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
