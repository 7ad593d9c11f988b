"""Builds the test harnesses against the HIP library in-tree:
tests/harness/ac_hermite (shared-step Hermite over the legacy ABI) and
tests/harness/ac_full (two-time-scale Ahmad-Cohen over the device-resident
entry points), which also link the oracle library as their reference
provider; and, linking the HIP library only, for bench.py:
tests/harness/ac_hermite_gpu (the same Hermite run without the oracle
provider) and tests/harness/legacy_timing (the legacy calling pattern timed
from C)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ac_hermite")
# binary -> (source, links the oracle library, extra flags)
NAMES = {"ac_hermite": ("ac_hermite", True, []),
         "ac_full": ("ac_full", True, []),
         "ac_hermite_gpu": ("ac_hermite", False, ["-DAC_NO_ORACLE"]),
         "legacy_timing": ("legacy_timing", False, [])}


def build(verbose=False):
    for name, (src, with_oracle, flags) in NAMES.items():
        out = os.path.join(HERE, name)
        oracle = ["-L", os.path.join(ROOT, "oracle"), "-loracle"] if with_oracle else []
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", *flags, "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, src + ".cpp"),
               "-L", os.path.join(ROOT, "abyss_amd"), "-labyss_regf", *oracle,
               "-Wl,-rpath,$ORIGIN/../../abyss_amd" + (":$ORIGIN/../../oracle" if with_oracle else ""),
               "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
