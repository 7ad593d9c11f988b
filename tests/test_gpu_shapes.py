"""GPU tests of the workgroup shapes of the lanes-per-target kernels (j per
LDS tile, wave64s per workgroup) and of the autotuner that picks them.

* Every shape gives the bits of the default shape (256 j, 4 waves) for the
  same lanes per target and j-split count, in split and in checked
  precision.  A full tile is a divisor of the 64-step FP32 run and split
  spans are whole 256-j blocks, so each lane folds its runs at the same j
  whatever the tile; the workgroup only groups targets.  The one-shot tile
  (1024 j, a whole split of at most 1024 j) likewise, with 16 or 64 lanes
  per target; plans it cannot hold fall back to 256.  (Checked blocks of at
  most 256 targets run in FP64 terms in one pass, which has one shape.)
* A batched ensemble runs each cluster with the tile of the installed table
  and gives the bits it gives without the table (the table used keeps the
  built-in lanes-per-target rule, so only the shape differs).
* abyss_autotune_jset tunes on the caller's own j-set, writes a six-column
  table that loads back, and its plans meet the oracle bars.
"""
import sys

import numpy as np
import pytest

import oracle
from oracle.compare import compare_lists, force_errors
from abyss_amd import regf, workloads
from tests.test_gpu_ensemble import make_clusters, random_targets
from tests.test_gpu_parity import _same_set_rows

pytestmark = pytest.mark.gpu

SHAPES = [(256, 4), (128, 4), (64, 4), (128, 2), (64, 2), (64, 1)]
CHECKED_DIRECT_MAX = 256   # checked blocks up to this size: FP64 terms, one pass


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def plummer():
    return workloads.plummer_kroupa(65536)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("prec", [regf.PREC_SPLIT, regf.PREC_CHECKED])
def test_shapes_give_the_same_bits(plummer, prec):
    js = plummer
    rng = np.random.default_rng(17)
    cases = {"random3000": rng.permutation(js.n)[:3000].astype(np.int32),
             "first300": np.arange(300, dtype=np.int32),
             "random40": rng.permutation(js.n)[:40].astype(np.int32)}
    with regf.Context(device=0, cap=js.cap, precision=prec) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        for name, t in cases.items():
            for jl in (4, 16, 64):
                for splits in (1, 7):
                    say(f"case: shapes prec={prec} targets={name} jl={jl} splits={splits}")
                    c.set_tuning(1, splits, 2, jl)
                    base = None
                    for tile, wpb in SHAPES:
                        c.set_shape(tile, wpb)
                        r = c.query(t)
                        p = c.plan()
                        if prec == regf.PREC_CHECKED and len(t) <= CHECKED_DIRECT_MAX:
                            assert (p["tile"], p["waves_per_block"]) == (256, 4), p
                        else:
                            assert (p["tile"], p["waves_per_block"]) == (tile, wpb), p
                            assert p["targets_per_block"] == wpb * 64 // jl, p
                        assert p["j_splits"] == splits, p
                        nre = c.last_recheck() if prec == regf.PREC_CHECKED else 0
                        if base is None:
                            base, base_nre = r, nre
                            continue
                        assert _same(r, base), (name, jl, splits, tile, wpb)
                        assert nre == base_nre
                    say(f"  {len(SHAPES)} shapes identical (recomputed {base_nre})")
        c.set_tuning()
        c.set_shape()
    # and the default shape meets the oracle bars on one case
    with regf.Context(device=0, cap=js.cap, precision=prec) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        t = cases["random3000"]
        c.set_shape(64, 1)
        r = c.query(t)
        ra, rj, rn, rl, sa, sj = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap, scales=True)
        err = force_errors(r[0], r[1], ra, rj, sa, sj)
        lists = compare_lists(r[2], r[3], rn, rl, js.x, js.r2, t)
        say(f"  (64, 1) vs oracle: a_norm_max={err['a_norm']['max']:.2e} lists={lists}")
        assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0
        assert err["a_norm"]["max"] <= 1e-6 and err["jdot_norm"]["max"] <= 1e-6


@pytest.mark.parametrize("prec", [regf.PREC_SPLIT, regf.PREC_CHECKED])
def test_one_shot_tile_gives_the_same_bits(plummer, prec):
    """The one-shot tile (each split's 1024 j loaded at once) against
    256-record tiles: the same bits, split precision and checked blocks of
    at most 256 targets (FP64 terms in one pass); built in below 16
    targets; plans it cannot hold fall back to 256 j."""
    js = plummer
    rng = np.random.default_rng(23)
    cases = {"random1": rng.permutation(js.n)[:1].astype(np.int32),
             "random16": rng.permutation(js.n)[:16].astype(np.int32),
             "first64": np.arange(64, dtype=np.int32),
             "random200": rng.permutation(js.n)[:200].astype(np.int32)}
    with regf.Context(device=0, cap=js.cap, precision=prec) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        for name, t in cases.items():
            for jl in (16, 64):
                say(f"case: one-shot tile prec={prec} targets={name} jl={jl}")
                c.set_tuning(1, 64, 2, jl)     # splits of 1024 j
                c.set_shape(256, 4)
                base = c.query(t)
                base_nre = c.last_recheck() if prec == regf.PREC_CHECKED else 0
                c.set_shape(1024, 4)
                r = c.query(t)
                p = c.plan()
                assert (p["tile"], p["j_splits"]) == (1024, 64), p
                assert _same(r, base), (name, jl)
                if prec == regf.PREC_CHECKED:
                    assert c.last_recheck() == base_nre
        # the built-in plan of a block below 16 targets uses it
        c.set_tuning()
        c.set_shape()
        c.query(cases["random1"])
        p = c.plan()
        assert (p["lanes_per_target"], p["tile"]) == (64, 1024), p
        # plans the one-shot tile cannot hold fall back to 256 j
        for jl, splits in ((4, 64), (16, 7), (64, 1)):
            c.set_tuning(1, splits, 2, jl)
            c.set_shape(1024, 0)
            c.query(cases["random16"])
            p = c.plan()
            assert p["tile"] == 256 and p["waves_per_block"] == 4, (jl, splits, p)
        c.set_tuning()
        c.set_shape()


def test_shapes_full_n_king():
    """The full-N King step (N = 262144, every particle a target) in three
    shapes, split precision: identical bits."""
    say("case: shapes on king_w6_binaries N=262144 full")
    js = workloads.king_binaries(262144)
    t = np.arange(js.n, dtype=np.int32)
    with regf.Context(device=0, cap=js.cap, precision=regf.PREC_SPLIT) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        c.set_tuning(1, 2, 2, 4)
        base = None
        for tile, wpb in ((256, 4), (128, 2), (64, 1)):
            c.set_shape(tile, wpb)
            r = c.query(t)
            say(f"  shape ({tile}, {wpb}): plan {c.plan()}")
            if base is None:
                base = r
            else:
                assert _same(r, base), (tile, wpb)


def test_shape_arguments_are_checked():
    with regf.Context(device=0, cap=64) as c:
        for tile, wpb in ((32, 4), (512, 4), (256, 2), (128, 1), (64, 3), (1024, 2)):
            with pytest.raises(regf.AbyssError):
                c.set_shape(tile, wpb)
        c.set_shape(256, 0)   # the other one automatic
        c.set_shape(0, 1)


def _write_table(path, tile):
    """A table with the built-in lanes-per-target rule (4 from 256 targets,
    16 from 16, 64 below) and one tile for every bucket."""
    with open(path, "w") as f:
        f.write("# bucket jl unroll wg_per_cu tile waves\n")
        for b in range(32):
            jl = 4 if b >= 8 else (16 if b >= 4 else 64)
            f.write(f"{b} {jl} 2 {48 if jl == 4 else 16} {tile} 4\n")


def test_ensemble_uses_table_tiles(tmp_path):
    sizes = [1024, 5000, 300, 20000, 65536, 2000]
    clusters = make_clusters(sizes, seed0=300)
    targets = [random_targets(len(cl[2]), 41 + k) for k, cl in enumerate(clusters)]
    results = {}
    try:
        for tile in (None, 128, 64):
            if tile is None:
                regf.autotune_clear(0)
            else:
                path = str(tmp_path / f"tile{tile}.txt")
                _write_table(path, tile)
                regf.autotune_load(0, path)
            for prec in (regf.PREC_SPLIT, regf.PREC_CHECKED):
                say(f"case: ensemble with table tile={tile} prec={prec}")
                with regf.Ensemble(device=0, cap=128, max_clusters=len(clusters)) as ens:
                    ens.set_precision(prec)
                    for k, (x, v, m, r2) in enumerate(clusters):
                        ens.upload(k, m, x, v, r2)
                        ens.set_targets(k, targets[k])
                    for k in range(len(clusters)):
                        jl, t_used = ens.plan(k)
                        assert t_used == (tile or 256), (k, jl, t_used)
                    ens.launch()
                    ens.sync()
                    results[(tile, prec)] = [ens.fetch(k) for k in range(len(clusters))]
    finally:
        regf.autotune_clear(0)
    for prec in (regf.PREC_SPLIT, regf.PREC_CHECKED):
        for tile in (128, 64):
            for k in range(len(clusters)):
                assert _same(results[(tile, prec)][k], results[(None, prec)][k]), (tile, prec, k)
        say(f"  prec={prec}: tiles 128 and 64 give the bits of the built-in plan")


def test_autotune_on_own_jset(tmp_path, plummer):
    js = plummer
    say("case: autotune on the caller's j-set (Plummer-Kroupa N=65536 subset of 16384)")
    sub = np.random.default_rng(5).permutation(js.n)[:16384]
    m, x, v, r2 = js.m[sub], js.x[sub], js.v[sub], js.r2[sub]
    path = str(tmp_path / "own.txt")
    try:
        regf.autotune_jset(0, m, x, v, r2, path)
        rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
        assert len(rows) == 32 and all(len(r) == 6 for r in rows)
        shapes = {(int(r[4]), int(r[5])) for r in rows}
        say(f"  tuned shapes: {sorted(shapes)}")
        assert shapes <= set(SHAPES) | {(1024, 4)}
        regf.autotune_clear(0)
        regf.autotune_load(0, path)
        with regf.Context(device=0, cap=js.cap) as c:
            c.upload(js.m, js.x, js.v, js.r2)
            for size in (1, 100, 5000):
                t = np.random.default_rng(size).permutation(js.n)[:size].astype(np.int32)
                r = c.query(t)
                say(f"  tuned plan for {size} targets: {c.plan()}")
                ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap)
                lists = compare_lists(r[2], r[3], rn, rl, js.x, js.r2, t)
                assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0
                # checked precision: 1e-6 relative on every row whose regular
                # set the FP64 evaluation agrees with
                rows = _same_set_rows(r[2], r[3], rn, rl)
                err = force_errors(r[0][rows], r[1][rows], ra[rows], rj[rows])
                assert err["a_rel"]["max"] <= 1e-6 and err["jdot_rel"]["max"] <= 1e-6, err
    finally:
        regf.autotune_clear(0)
