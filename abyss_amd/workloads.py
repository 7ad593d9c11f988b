"""The regular-force workloads, built from the seeded IC generators.

plummer_equal(n=1024)   equal-mass Plummer, seed 1, ~48 neighbours
plummer_kroupa(n=65536) Plummer with Kroupa IMF, ~100 neighbours, cap 256
king_binaries(n=262144) King W0=6, Kroupa IMF, 5% of particles in hard
                        binaries (a ~ log-U[1e-5, 1e-3]), ~100 neighbours
ensemble_clusters()     256 Plummer clusters, N log-uniform in [1024, 65536],
                        active fraction uniform in [0.01, 1] (batched ensemble)
served_clusters()       64 clusters, N log-uniform in [4096, 1048576], Plummer
                        or King W0 uniform in [3, 9] (served ensemble)
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


def ensemble_clusters(count=256, lo=1024, hi=65536, seed=1, k_nb=64):
    """`count` Plummer clusters, N log-uniform in [lo, hi], distinct seeds,
    active fraction uniform in [0.01, 1]; Kroupa masses; r_nb for ~k_nb
    neighbours.  Returns [(x, v, m, r2, targets)]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(int)
    fr = rng.uniform(0.01, 1.0, count)
    out = []
    for k in range(count):
        n = int(ns[k])
        r = np.random.default_rng(seed * 1000 + k)
        x, v, m = ic.plummer(n, seed=seed * 1000 + k, masses=ic.kroupa_masses(n, r))
        r2 = ic.neighbour_radius2(x, k_nb)
        t = r.permutation(n)[:max(1, int(round(fr[k] * n)))].astype(np.int32)
        out.append((x, v, m, r2, t))
    return out


def served_clusters(count=64, lo=4096, hi=1048576, seed=5, k_nb=32, sizes=None):
    """`count` clusters, N log-uniform in [lo, hi] (or the given `sizes`),
    Plummer (even k) or King with W0 uniform in [3, 9] (odd k), Kroupa
    masses, r_nb for ~k_nb neighbours.  Returns [(x, v, m, r2)]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(int)
    if sizes is not None:
        ns = np.asarray(sizes, dtype=int)
    out = []
    for k in range(len(ns)):
        n = int(ns[k])
        r = np.random.default_rng(seed * 1000 + k)
        masses = ic.kroupa_masses(n, r)
        if k % 2:
            x, v, m = ic.king(n, W0=float(r.uniform(3.0, 9.0)), seed=seed * 1000 + k,
                              masses=masses)
        else:
            x, v, m = ic.plummer(n, seed=seed * 1000 + k, masses=masses)
        out.append((x, v, m, ic.neighbour_radius2(x, k_nb)))
    return out
