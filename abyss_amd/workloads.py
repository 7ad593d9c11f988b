"""The regular-force workloads, built from the seeded IC generators.

plummer_equal(n=1024)   equal-mass Plummer, seed 1, ~48 neighbours
plummer_kroupa(n=65536) Plummer with Kroupa IMF, ~100 neighbours, cap 256
king_binaries(n=262144) King W0=6, Kroupa IMF, 5% of particles in hard
                        binaries (a ~ log-U[1e-5, 1e-3]), ~100 neighbours
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import ic


@dataclass
class JSet:
    name: str
    x: np.ndarray
    v: np.ndarray
    m: np.ndarray
    r2: np.ndarray          # r_nb^2 per particle
    cap: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.m)


_cache = {}


def plummer_equal(n=1024, seed=1, k_nb=48, cap=256):
    key = ("pe", n, seed, k_nb)
    if key not in _cache:
        x, v, m = ic.plummer(n, seed=seed)
        r2 = ic.neighbour_radius2(x, k_nb)
        _cache[key] = JSet(f"plummer_equal_n{n}", x, v, m, r2, cap,
                           {"seed": seed, "k_nb": k_nb})
    return _cache[key]


def plummer_kroupa(n=65536, seed=1, k_nb=100, cap=256):
    key = ("pk", n, seed, k_nb)
    if key not in _cache:
        rng = np.random.default_rng(seed)
        m = ic.kroupa_masses(n, rng)
        x, v, m = ic.plummer(n, seed=seed, masses=m)
        r2 = ic.neighbour_radius2(x, k_nb)
        _cache[key] = JSet(f"plummer_kroupa_n{n}", x, v, m, r2, cap,
                           {"seed": seed, "k_nb": k_nb})
    return _cache[key]


def king_binaries(n=262144, W0=6.0, seed=1, k_nb=100, cap=256, bin_frac=0.05):
    key = ("kb", n, W0, seed, k_nb, bin_frac)
    if key not in _cache:
        rng = np.random.default_rng(seed)
        m = ic.kroupa_masses(n, rng)
        x, v, m = ic.king(n, W0=W0, seed=seed, masses=m)
        x, v, m, pairs = ic.add_hard_binaries(x, v, m, frac=bin_frac, seed=seed)
        r2 = ic.neighbour_radius2(x, k_nb)
        _cache[key] = JSet(f"king_w{W0:g}_binaries_n{n}", x, v, m, r2, cap,
                           {"seed": seed, "k_nb": k_nb, "W0": W0, "pairs": pairs})
    return _cache[key]


def active_blocks(n, count, seed=1, lo=1, hi=None):
    """`count` target lists with sizes log-uniform on [lo, hi], random order."""
    hi = n if hi is None else hi
    sizes = ic.log_uniform_block_sizes(count, lo, hi, seed=seed)
    rng = np.random.default_rng(seed + 101)
    return [rng.permutation(n)[:s].astype(np.int32) for s in sizes]
