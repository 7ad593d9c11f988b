"""Writes the golden regular-force fixtures (inputs + FP64 oracle outputs).

The reference ships no fixtures for this path, so these are produced by the
pinned oracle (oracle/regf_oracle.c, checked in tests/test_oracle.py) on
seeded synthetic clusters:
  plummer_n64_seed1.npz    equal-mass Plummer, N=64, r_nb for ~8 neighbours
  plummer_n1024_seed1.npz  equal-mass Plummer, N=1024, r_nb for ~48 neighbours
Each file holds x, v, m, r2 (inputs), targets, cap and acc, adot, nnb, lists
(outputs).  Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from abyss_amd import ic  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def make(n, k_nb, cap, name):
    x, v, m = ic.plummer(n, seed=1)
    r2 = ic.neighbour_radius2(x, k_nb)
    t = np.arange(n, dtype=np.int32)
    acc, adot, nnb, lists = oracle.regf(x, v, m, r2, t, cap)
    np.savez_compressed(os.path.join(HERE, name), x=x, v=v, m=m, r2=r2, targets=t,
                        cap=np.int32(cap), acc=acc, adot=adot, nnb=nnb, lists=lists)


if __name__ == "__main__":
    make(64, 8, 32, "plummer_n64_seed1.npz")
    make(1024, 48, 96, "plummer_n1024_seed1.npz")
    print("written")
