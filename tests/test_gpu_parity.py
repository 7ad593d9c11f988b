"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

Bars:
  * neighbour lists (every mode): ascending, identical to the oracle's except
    pairs with |r^2 - r_nb^2| <= 1e-6 r_nb^2, which are counted
    (hard_mismatches == 0);
  * the default precision (checked), through contexts and through the six
    legacy entry points: a and jdot within 1e-6 relative of FP64 on EVERY
    target, |d a| / |a| <= 1e-6 at the maximum (rows whose regular set the
    FP64 evaluation agrees with; a row that differs by a band pair sums a
    different set and is compared through its list);
  * split precision (opt-in, FP32 terms without the check): relative error
    <= 1e-6 at the 99th percentile and normalised error |d a| /
    sum_j |term_j| <= 1e-6 at the maximum (a particle whose sum nearly
    cancels can exceed 1e-6 relative under FP32 terms);
  * FP64 kernel: relative error <= 1e-12 and lists equal outside a 1e-12 band.
"""
import sys

import numpy as np
import pytest

import oracle
from oracle.compare import compare_lists, force_errors
from abyss_amd import regf, workloads

pytestmark = pytest.mark.gpu

TOL_REL_P99 = 1e-6
TOL_NORM_MAX = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    """Split precision (the split-specific tests restore it after a switch)."""
    c = regf.Context(device=0, cap=256, precision=regf.PREC_SPLIT)
    yield c
    c.close()


@pytest.fixture(scope="module")
def dctx():
    """The library's default precision."""
    c = regf.Context(device=0, cap=256)
    yield c
    c.close()


def check(js, targets, res, band=1e-6, prec=regf.PREC_SPLIT, label=""):
    acc, adot, nnb, lst = res
    ra, rj, rn, rl, sa, sj = oracle.regf(js.x, js.v, js.m, js.r2, targets, js.cap, scales=True)
    err = force_errors(acc, adot, ra, rj, sa, sj)
    lists = compare_lists(nnb, lst, rn, rl, js.x, js.r2, targets, band=band)
    say(f"[{label}] n={js.n} nt={len(targets)} a_rel={err['a_rel']} jdot_rel={err['jdot_rel']} "
        f"a_norm_max={err['a_norm']['max']:.3e} jdot_norm_max={err['jdot_norm']['max']:.3e} "
        f"lists={lists}")
    assert lists["hard_mismatches"] == 0
    assert lists["not_ascending"] == 0
    assert lists["overflow_mismatch"] == 0
    if prec == regf.PREC_FP64:
        assert err["a_rel"]["max"] <= 1e-12
        assert err["jdot_rel"]["max"] <= 1e-12
    elif prec == regf.PREC_REFINED:
        # FP64 terms with a Newton-refined rsqrt: FP64-level per term
        assert err["a_rel"]["p99"] <= 1e-12 and err["jdot_rel"]["p99"] <= 1e-12
        assert err["a_norm"]["max"] <= 1e-12 and err["jdot_norm"]["max"] <= 1e-12
    else:
        assert err["a_rel"]["p99"] <= TOL_REL_P99
        assert err["jdot_rel"]["p99"] <= TOL_REL_P99
        assert err["a_norm"]["max"] <= TOL_NORM_MAX
        assert err["jdot_norm"]["max"] <= TOL_NORM_MAX
    return err, lists


def test_two_body(ctx):
    say("case: two_body")
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    v = np.array([[0.1, -0.2, 0.3], [0.4, 0.1, -0.5]])
    m = np.array([1.5, 2.5])
    ctx.upload(m, x, v, np.zeros(2))
    acc, adot, nnb, _ = ctx.query([0, 1])
    np.testing.assert_allclose(acc[0], [0.06, 0.08, 0.0], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(adot[0], [-0.00912, -0.01416, -0.016], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(acc[1], [-0.036, -0.048, 0.0], rtol=1e-6, atol=1e-9)
    assert list(nnb) == [0, 0]


def test_plummer_1024_full(ctx):
    say("case: plummer_equal N=1024 full")
    js = workloads.plummer_equal(1024)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.arange(js.n, dtype=np.int32)
    check(js, t, ctx.query(t), label="plummer1024")


@pytest.mark.parametrize("size", [1, 3, 64, 1000, 8192, 65536])
def test_plummer_kroupa_65536_blocks(ctx, size):
    say(f"case: plummer_kroupa N=65536 block={size}")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    rng = np.random.default_rng(size)
    t = rng.permutation(js.n)[:size].astype(np.int32)
    err, _ = check(js, t, ctx.query(t), label=f"kroupa65536/{size}")


def test_switching_block_sizes_on_one_context(ctx):
    """Small blocks (<= 64 targets) take the mapped target list and the
    finalize's host copy; larger ones the device copies.  Alternating sizes
    on one context, through query, launch + fetch and preallocated outputs,
    must give every block its own correct results."""
    say("case: switching block sizes on one context")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    rng = np.random.default_rng(77)
    first = None
    for k, size in enumerate([1, 64, 65, 3, 300, 2, 64, 5000, 7, 1]):
        t = rng.permutation(js.n)[:size].astype(np.int32)
        if k % 3 == 0:
            res = ctx.query(t)
        elif k % 3 == 1:
            ctx.set_targets(t)
            ctx.launch()
            ctx.launch()               # a second launch rewrites the same results
            res = ctx.fetch()
        else:
            out = (np.zeros((size, 3)), np.zeros((size, 3)), np.zeros(size, np.int32),
                   np.zeros((size, ctx.cap), np.int32))
            res = ctx.query(t, out=out)
            assert res[0] is out[0] and res[3] is out[3]
        check(js, t, res, label=f"switch/{k}/{size}")
        if first is None:
            first = (t, [np.copy(a) for a in res])
    # the first block again: bitwise what it gave the first time
    again = ctx.query(first[0])
    for a, b in zip(first[1], again):
        assert np.array_equal(a, b)


def test_fp64_kernel(ctx):
    say("case: fp64 kernel plummer_kroupa N=65536 block=4096")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(5).permutation(js.n)[:4096].astype(np.int32)
    ctx.set_precision(regf.PREC_FP64)
    try:
        check(js, t, ctx.query(t), band=1e-12, prec=regf.PREC_FP64, label="fp64")
    finally:
        ctx.set_precision(regf.PREC_SPLIT)


@pytest.mark.parametrize("prec", [regf.PREC_MIXED, regf.PREC_REFINED])
@pytest.mark.parametrize("size", [1, 100, 4096, 65536])
def test_other_precision_modes(ctx, prec, size):
    name = {regf.PREC_MIXED: "mixed", regf.PREC_REFINED: "refined"}[prec]
    say(f"case: {name} precision plummer_kroupa N=65536 block={size}")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(size + 17).permutation(js.n)[:size].astype(np.int32)
    ctx.set_precision(prec)
    try:
        # these modes test neighbours in FP64: lists exact to 1e-12
        check(js, t, ctx.query(t), band=1e-12, prec=prec, label=f"{name}/{size}")
        assert ctx.plan()["lanes_per_target"] > 1
    finally:
        ctx.set_precision(regf.PREC_SPLIT)


def _same_set_rows(nnb, lst, rn, rl):
    """Rows whose neighbour set equals the reference's (a pair in the 1e-6
    band may fall on the other side in FP64; such a row sums a different
    regular set and is compared through its list instead)."""
    w = lst.shape[1]
    col = np.arange(w)[None, :]
    valid = col < np.minimum(np.abs(nnb), w)[:, None]
    return (nnb == rn) & np.all((lst == rl[:, :w]) | ~valid, axis=1)


CHECK_TOL = 1e-6
# checked blocks of at most this many targets run in FP64 terms in one pass
CHECKED_DIRECT_MAX = 256


@pytest.mark.parametrize("wl,size", [("plummer1024", 1024), ("kroupa", 1), ("kroupa", 100),
                                     ("kroupa", 4096), ("kroupa", 65536)])
def test_checked_precision(ctx, wl, size):
    """Checked precision: every result within 1e-6 relative of the FP64
    oracle (a and jdot, per particle, maximum) and lists and counts those of
    split precision.  Blocks above 256 targets: unflagged targets bitwise
    split precision, recomputed targets FP64-accurate, and the recomputed
    set the one the flag rule names.  Smaller blocks: every target in FP64
    terms in one pass, bitwise refined precision where its neighbour test
    agrees."""
    say(f"case: checked precision {wl} block={size}")
    js = workloads.plummer_equal(1024) if wl == "plummer1024" else workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(size + 29).permutation(js.n)[:size].astype(np.int32)
    split = ctx.query(t)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        chk = ctx.query(t)
        nre = ctx.last_recheck()
        if size <= CHECKED_DIRECT_MAX:
            with pytest.raises(regf.AbyssError):
                ctx.spread()
            ctx.set_precision(regf.PREC_REFINED)
            refined = ctx.query(t)
        else:
            spr = ctx.spread().astype(np.float64)
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    assert np.array_equal(chk[2], split[2]) and np.array_equal(chk[3], split[3])
    ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap)
    rows = _same_set_rows(chk[2], chk[3], rn, rl)
    err = force_errors(chk[0][rows], chk[1][rows], ra[rows], rj[rows])
    say(f"[checked {wl}/{size}] recomputed={nre} a_rel={err['a_rel']['max']:.2e} "
        f"jdot_rel={err['jdot_rel']['max']:.2e}")
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL
    if size <= CHECKED_DIRECT_MAX:
        # a small block: every target in FP64 terms in one pass, the sums of
        # refined precision wherever the two neighbour tests agree (the
        # refined test compares the FP64 r^2; rows with a pair in that band
        # sum a different regular set)
        assert nre == size
        same = np.all(refined[3] == split[3], axis=1) & (refined[2] == split[2])
        assert same.sum() >= size - 2
        assert np.array_equal(chk[0][same], refined[0][same])
        assert np.array_equal(chk[1][same], refined[1][same])
        return
    changed = np.any(chk[0] != split[0], axis=1) | np.any(chk[1] != split[1], axis=1)
    assert changed.sum() <= nre <= size
    # the flag rule, evaluated on the host from the split results (a 1 %
    # margin each way for the FP32 spread's rounding)
    na = np.linalg.norm(split[0], axis=1)
    nj = np.linalg.norm(split[1], axis=1)
    score = np.maximum(spr[:, 0] / np.where(na > 0, na, 1e-300),
                       spr[:, 1] / np.where(nj > 0, nj, 1e-300)) * regf.CHECK_ERR_SCALE / CHECK_TOL
    assert np.all(changed[score > 1.01]) and not np.any(changed[score < 0.99])
    ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap)
    rows = _same_set_rows(chk[2], chk[3], rn, rl)
    err = force_errors(chk[0][rows], chk[1][rows], ra[rows], rj[rows])
    sp_err = force_errors(split[0][rows], split[1][rows], ra[rows], rj[rows])
    say(f"[checked {wl}/{size}] recomputed={nre} band_rows={(~rows).sum()} "
        f"a_rel={err['a_rel']} jdot_rel={err['jdot_rel']} (split: jdot_rel={sp_err['jdot_rel']})")
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL
    if changed[rows].any():
        rerr = force_errors(chk[0][rows & changed], chk[1][rows & changed], ra[rows & changed],
                            rj[rows & changed])
        assert rerr["a_rel"]["max"] <= 1e-11 and rerr["jdot_rel"]["max"] <= 1e-11


@pytest.mark.parametrize("size", [8, 64, 256])
def test_checked_precision_small_block_in_fp64_terms(ctx, size):
    """Blocks of up to 256 targets evaluate every target in FP64 terms in one
    pass: no spread, every target counted as recomputed, lists those of
    split precision, and FP64 accuracy on the targets that a full-N checked
    launch had to recompute as on the others."""
    say(f"case: checked precision, small block of {size} flagged-prone targets")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    full = np.arange(js.n, dtype=np.int32)
    split_full = ctx.query(full)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        chk_full = ctx.query(full)
        hot = np.flatnonzero(np.any(chk_full[1] != split_full[1], axis=1))
        cold = np.setdiff1d(full, hot)
        t = np.concatenate([hot[: size * 3 // 4], cold[: size - min(len(hot), size * 3 // 4)]])
        t = np.random.default_rng(size).permutation(t).astype(np.int32)
        chk = ctx.query(t)
        nre = ctx.last_recheck()
        with pytest.raises(regf.AbyssError):
            ctx.spread()
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    split = ctx.query(t)
    assert len(t) == size and ctx.plan()["j_splits"] > 1
    say(f"[checked small {size}] recomputed={nre}")
    assert nre == size
    assert np.array_equal(chk[2], split[2]) and np.array_equal(chk[3], split[3])
    ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap)
    rows = _same_set_rows(chk[2], chk[3], rn, rl)
    err = force_errors(chk[0][rows], chk[1][rows], ra[rows], rj[rows])
    say(f"  a_rel={err['a_rel']['max']:.2e} jdot_rel={err['jdot_rel']['max']:.2e}")
    assert err["a_rel"]["max"] <= 1e-11 and err["jdot_rel"]["max"] <= 1e-11
    # the legacy entry points (their own context, list stride 1000) give the
    # same results for the same block
    regf.OpenDevice(0)
    try:
        regf.set_precision(regf.PREC_CHECKED)
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        acc, adot, nnb, _ = regf.CalculateAccelerationOnDevice(t)
    finally:
        regf.set_precision(regf.PREC_SPLIT)
        regf.CloseDevice()
    assert np.array_equal(acc, chk[0]) and np.array_equal(adot, chk[1])
    assert np.array_equal(nnb, chk[2])


def test_checked_count_read_on_device(ctx):
    """Checked blocks of 257..4096 targets read the flagged count on the
    device: the launch queues the recomputation without a host wait, and
    last_recheck waits for the reported count.  Two launches in a row with
    no sync, then the count and the results equal a synchronous query's."""
    say("case: checked count read on the device")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(41).permutation(js.n)[:2048].astype(np.int32)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        ref = ctx.query(t)
        nref = ctx.last_recheck()
        ctx.set_targets(t)
        ctx.launch()
        ctx.launch()
        n2 = ctx.last_recheck()
        out = ctx.fetch()
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    say(f"  recomputed {nref} / {n2} of {len(t)}")
    assert 0 < nref < len(t) and n2 == nref
    for a, b in zip(out, ref):
        assert np.array_equal(a, b)


def test_checked_precision_tolerance_and_legacy(ctx):
    """A tighter tolerance recomputes more targets; the legacy entry points
    give the context's checked results bit for bit."""
    say("case: checked precision tolerance + legacy")
    js = workloads.plummer_equal(1024)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.arange(js.n, dtype=np.int32)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        ref = ctx.query(t)
        n_default = ctx.last_recheck()
        # deterministic: the flagged order varies (atomic appends), but each
        # target's recomputation does not depend on its place in that order
        again = ctx.query(t)
        assert ctx.last_recheck() == n_default
        for a, b in zip(ref, again):
            assert np.array_equal(a, b)
        ctx.set_check_tolerance(1e-7)
        ctx.query(t)
        n_tight = ctx.last_recheck()
        ctx.set_check_tolerance(1.0)
        loose = ctx.query(t)
        assert ctx.last_recheck() == 0
        with pytest.raises(regf.AbyssError):
            ctx.set_check_tolerance(0.0)
        ctx.set_check_tolerance(CHECK_TOL)
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    split = ctx.query(t)
    say(f"[checked tolerance] recomputed default={n_default} tight={n_tight}")
    assert n_tight > n_default > 0
    assert np.array_equal(loose[0], split[0]) and np.array_equal(loose[1], split[1])
    regf.OpenDevice(0)
    try:
        regf.set_precision(regf.PREC_CHECKED)
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        acc, adot, nnb, lst = regf.CalculateAccelerationOnDevice(t)
    finally:
        regf.set_precision(regf.PREC_SPLIT)
        regf.CloseDevice()
    assert np.array_equal(acc, ref[0]) and np.array_equal(adot, ref[1])
    assert np.array_equal(nnb, ref[2])


def test_tuning_variants_agree(ctx):
    """Every kernel variant classifies every pair identically (same FP32 r^2),
    so lists are equal; one-target-per-lane variants share FP32 partial groups
    (folds aligned to absolute j) and differ only in FP64 summation order;
    lanes-per-target variants group the FP32 partials differently."""
    say("case: tuning variants")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(9).permutation(js.n)[:3000].astype(np.int32)
    base = None
    try:
        variants = [(ipl, s_, un, 1) for ipl in (1, 2, 4) for s_ in (1, 5, 0) for un in (2, 4)]
        variants += [(1, s_, 2, jl) for jl in (4, 16, 64) for s_ in (1, 7, 0)]
        for ipl, splits, un, jl in variants:
            ctx.set_tuning(ipl, splits, un, jl)
            r = ctx.query(t)
            say(f"  ipl={ipl} splits={splits} unroll={un} jl={jl} plan={ctx.plan()}")
            if base is None:
                base = r
            if jl == 1:
                np.testing.assert_allclose(r[0], base[0], rtol=1e-10,
                                           atol=1e-12 * np.abs(base[0]).max())
                np.testing.assert_allclose(r[1], base[1], rtol=1e-10,
                                           atol=1e-12 * np.abs(base[1]).max())
            else:
                check(js, t, r, label=f"jl{jl}/S{splits}")
            assert np.array_equal(r[2], base[2])
            assert np.array_equal(r[3], base[3])
    finally:
        ctx.set_tuning(0, 0, 0, 0)


def test_deterministic(ctx):
    say("case: determinism")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(3).permutation(js.n)[:20000].astype(np.int32)
    r1 = ctx.query(t)
    r2 = ctx.query(t)
    for a, b in zip(r1, r2):
        assert np.array_equal(a, b)


def test_overflow(ctx):
    say("case: overflow")
    js = workloads.plummer_equal(1024)
    small = regf.Context(device=0, cap=16)
    try:
        small.upload(js.m, js.x, js.v, js.r2)
        t = np.arange(0, js.n, 7, dtype=np.int32)
        acc, adot, nnb, lst = small.query(t)
        assert small.last_status == regf.ERR_NEIGHBOR_OVERFLOW
        ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, 16)
        assert np.array_equal(nnb, rn)
        assert np.all(nnb < 0)
        assert np.array_equal(lst, rl)
    finally:
        small.close()


@pytest.mark.parametrize("tuning", [(0, 0, 0, 16), (0, 37, 0, 4), (0, 3, 0, 64)])
def test_checked_precision_forced_plans(ctx, tuning):
    """Forced lane groups and split counts: the flag and the recomputation
    follow the plan of the split pass, and every result stays within 1e-6
    of the oracle."""
    say(f"case: checked precision with forced tuning {tuning}")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(71).permutation(js.n)[:2048].astype(np.int32)
    ctx.set_tuning(*tuning)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        chk = ctx.query(t)
        nre = ctx.last_recheck()
        plan = ctx.plan()
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
        ctx.set_tuning()
    assert plan["lanes_per_target"] == tuning[3]
    if tuning[1]:
        assert plan["j_splits"] == tuning[1]
    ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap)
    rows = _same_set_rows(chk[2], chk[3], rn, rl)
    err = force_errors(chk[0][rows], chk[1][rows], ra[rows], rj[rows])
    say(f"[checked forced {tuning}] recomputed={nre} a_rel={err['a_rel']['max']:.2e} "
        f"jdot_rel={err['jdot_rel']['max']:.2e}")
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL


def test_checked_precision_on_predicted_positions(ctx):
    """Checked precision on a j-set predicted on the device."""
    say("case: checked precision after device prediction, plummer_kroupa N=65536")
    js = workloads.plummer_kroupa(65536)
    rng = np.random.default_rng(13)
    a = rng.normal(size=(js.n, 3)) * 0.1
    ad = rng.normal(size=(js.n, 3)) * 0.1
    t0 = rng.uniform(0.0, 0.01, js.n)
    ctx.upload_state(js.m, js.x, js.v, a, ad, t0, js.r2)
    ctx.predict(0.01)
    t = rng.permutation(js.n)[:4096].astype(np.int32)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        chk = ctx.query(t)
        nre = ctx.last_recheck()
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    xp, vp = oracle.predict2(js.x, js.v, a, ad, 0.01 - t0)
    ra, rj, rn, rl = oracle.regf(xp, vp, js.m, js.r2, t, js.cap)
    rows = _same_set_rows(chk[2], chk[3], rn, rl)
    err = force_errors(chk[0][rows], chk[1][rows], ra[rows], rj[rows])
    say(f"[checked predicted] recomputed={nre} band_rows={(~rows).sum()} "
        f"a_rel={err['a_rel']['max']:.2e} jdot_rel={err['jdot_rel']['max']:.2e}")
    assert nre > 0 and (~rows).sum() <= 8
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL


def test_checked_precision_overflowed_lists():
    """Overflowed lists (cap 16, ~48 neighbours each): the recomputation
    still sums exactly the regular set (its test does not read the lists),
    so every result is within 1e-6 of the oracle and the truncated lists
    and negative counts are those of split precision."""
    say("case: checked precision, overflowed lists")
    js = workloads.plummer_equal(1024)
    small = regf.Context(device=0, cap=16, precision=regf.PREC_SPLIT)
    try:
        small.upload(js.m, js.x, js.v, js.r2)
        t = np.arange(js.n, dtype=np.int32)
        split = small.query(t)
        small.set_precision(regf.PREC_CHECKED)
        chk = small.query(t)
        assert small.last_status == regf.ERR_NEIGHBOR_OVERFLOW
        nre = small.last_recheck()
    finally:
        small.close()
    assert nre > 0
    assert np.array_equal(chk[2], split[2]) and np.all(chk[2] < 0)
    assert np.array_equal(chk[3], split[3])
    ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, 16)
    assert np.array_equal(chk[2], rn)
    err = force_errors(chk[0], chk[1], ra, rj)
    say(f"[checked overflow] recomputed={nre} a_rel={err['a_rel']} jdot_rel={err['jdot_rel']}")
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL


def test_edge_cases(ctx):
    say("case: edge cases")
    # single particle: zero force, no neighbours
    ctx.upload([1.0], [[0.1, 0.2, 0.3]], [[0.0, 0.0, 0.0]], [0.5])
    acc, adot, nnb, _ = ctx.query([0])
    assert np.all(acc == 0) and np.all(adot == 0) and nnb[0] == 0
    # coincident particles become neighbours, no NaN
    x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [4.0, 2.0, 3.0], [9.0, 9.0, 9.0]])
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
    m = np.array([1.0, 2.0, 3.0, 0.0])
    ctx.upload(m, x, v, np.zeros(4))
    acc, adot, nnb, lst = ctx.query([0, 1, 1, 0])   # duplicates allowed (NumTarget <= N)
    assert np.all(np.isfinite(acc)) and np.all(np.isfinite(adot))
    assert list(nnb) == [1, 1, 1, 1]
    assert list(lst[:, 0]) == [1, 0, 0, 1]
    np.testing.assert_allclose(acc[0], [1.0 / 3.0, 0, 0], rtol=1e-6)
    # r_nb = 0 -> empty lists; r_nb = inf -> all neighbours, zero force
    rng = np.random.default_rng(1)
    for n in (257, 1000, 1023):
        x = rng.normal(size=(n, 3))
        v = rng.normal(size=(n, 3))
        m = rng.uniform(0.5, 1.5, n) / n
        t = rng.permutation(n)[:100].astype(np.int32)
        ctx.upload(m, x, v, np.zeros(n))
        acc, adot, nnb, _ = ctx.query(t)
        ra, rj, _, _ = oracle.regf(x, v, m, np.zeros(n), t, 1)
        assert np.all(nnb == 0)
        np.testing.assert_allclose(acc, ra, rtol=2e-5)
        big = regf.Context(device=0, cap=n)
        try:
            big.upload(m, x, v, np.full(n, np.inf))
            acc, adot, nnb, lst = big.query(t)
            assert np.all(acc == 0) and np.all(adot == 0)
            assert np.all(nnb == n - 1)
            for q, i in enumerate(t):
                assert list(lst[q, :n - 1]) == [j for j in range(n) if j != i]
        finally:
            big.close()


def test_invalid_inputs(ctx):
    say("case: invalid inputs")
    js = workloads.plummer_equal(1024)
    ctx.upload(js.m, js.x, js.v, js.r2)
    with pytest.raises(regf.AbyssError) as e:
        ctx.query([0, js.n])
    assert e.value.code == regf.ERR_TARGET_RANGE
    x = js.x.copy()
    x[5, 1] = np.nan
    with pytest.raises(regf.AbyssError) as e:
        ctx.upload(js.m, x, js.v, js.r2)
    assert e.value.code == regf.ERR_INVALID_INPUT
    # a rejected upload leaves no j-set: no query runs on stale positions
    with pytest.raises(regf.AbyssError) as e:
        ctx.query([1, 2])
    assert e.value.code == regf.ERR_NO_JSET
    ctx.upload(js.m, js.x, js.v, js.r2)
    ctx.set_targets(np.arange(8, dtype=np.int32))
    r2 = js.r2.copy()
    r2[3] = -1.0
    with pytest.raises(regf.AbyssError) as e:
        ctx.upload(js.m, js.x, js.v, r2)
    assert e.value.code == regf.ERR_INVALID_INPUT
    with pytest.raises(regf.AbyssError) as e:
        ctx.launch()      # the targets went with the j-set
    with pytest.raises(regf.AbyssError) as e:
        ctx.query([1, 2])
    assert e.value.code == regf.ERR_NO_JSET
    # a good upload restores everything
    ctx.upload(js.m, js.x, js.v, js.r2)
    acc, _, _, _ = ctx.query([1, 2])
    assert np.all(np.isfinite(acc))


def test_legacy_entry_points():
    say("case: legacy entry points")
    js = workloads.plummer_equal(1024)
    regf.OpenDevice(0)
    try:
        assert regf.num_neighbor_max() == 1000
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        t = np.arange(0, js.n, 3, dtype=np.int32)
        acc, adot, nnb, lst = regf.CalculateAccelerationOnDevice(t)
        ra, rj, rn, rl, sa, sj = oracle.regf(js.x, js.v, js.m, js.r2, t, 1000, scales=True)
        res = compare_lists(nnb, lst, rn, rl, js.x, js.r2, t)
        assert res["hard_mismatches"] == 0
        # the bars of check(): component-wise tolerances fail on components
        # that cancel, so compare per particle
        # the default (checked) precision: every target within 1e-6
        err = force_errors(acc, adot, ra, rj, sa, sj)
        assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL
        regf.ProfileDevice(0)
        # a second open only reports it
        regf.OpenDevice(0)
    finally:
        regf.CloseDevice()
    with pytest.raises(regf.AbyssError) as e:
        regf.CalculateAccelerationOnDevice([0])
    assert e.value.code == regf.ERR_NOT_OPEN


def test_king_262144_full_properties(ctx):
    """Full-size King W0=6 + hard binaries: split kernel vs the FP64 kernel on
    every target, and both against the oracle on a sampled subset."""
    say("case: king_w6_binaries N=262144 full")
    js = workloads.king_binaries(262144)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.arange(js.n, dtype=np.int32)
    split = ctx.query(t)
    ctx.set_precision(regf.PREC_FP64)
    try:
        f64 = ctx.query(t)
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    lists = compare_lists(split[2], split[3], f64[2], f64[3], js.x, js.r2, t)
    err = force_errors(split[0], split[1], f64[0], f64[1])
    say(f"[king262144 split vs fp64] a_rel={err['a_rel']} jdot_rel={err['jdot_rel']} lists={lists}")
    assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0
    assert err["a_rel"]["p99"] <= TOL_REL_P99 and err["jdot_rel"]["p99"] <= TOL_REL_P99
    # every binary member sees its companion as a neighbour
    pairs = js.meta["pairs"]
    for p, s in pairs[:200]:
        assert s in set(split[3][p, :split[2][p]].tolist())
    # oracle on a sample
    sample = np.random.default_rng(4).permutation(js.n)[:1024].astype(np.int32)
    sub = tuple(a[sample] for a in split)
    check(js, sample, sub, label="king262144/sample")


def test_king_262144_exact_scalings(ctx):
    """Size-independent properties of the full King step: doubling every mass
    doubles a and jdot exactly, doubling every velocity doubles jdot exactly
    and leaves a unchanged, and neither changes a list.  A power-of-two
    factor commutes with every rounding of the split arithmetic (FP64
    differences, FP32 terms and runs, FP64 folds and split sums), so the
    equalities are bit for bit; a dropped or misplaced mass or velocity
    factor in the kernel breaks them."""
    say("case: king_w6_binaries N=262144 exact mass / velocity scalings")
    js = workloads.king_binaries(262144)
    t = np.arange(js.n, dtype=np.int32)
    ctx.upload(js.m, js.x, js.v, js.r2)
    base = ctx.query(t)
    ctx.upload(2.0 * js.m, js.x, js.v, js.r2)
    m2 = ctx.query(t)
    ctx.upload(js.m, js.x, 2.0 * js.v, js.r2)
    v2 = ctx.query(t)
    for r in (m2, v2):
        assert np.array_equal(r[2], base[2]) and np.array_equal(r[3], base[3])
    assert np.array_equal(m2[0], 2.0 * base[0]) and np.array_equal(m2[1], 2.0 * base[1])
    assert np.array_equal(v2[0], base[0]) and np.array_equal(v2[1], 2.0 * base[1])


def test_permuted_jset_relabels_lists():
    """Full-N query of a permuted j-set (Plummer + Kroupa N = 65536): every
    list is the original one relabelled (and ascending in the new labels),
    counts are equal, and a / jdot agree to FP32 rounding (the summation
    order changes with the permutation)."""
    say("case: permuted j-set, plummer_kroupa N=65536 full")
    js = workloads.plummer_kroupa(65536)
    perm = np.random.default_rng(23).permutation(js.n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(js.n)
    with regf.Context(device=0, cap=js.cap, precision=regf.PREC_SPLIT) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        a0, j0, n0, l0 = c.query(np.arange(js.n, dtype=np.int32))
        c.upload(js.m[perm], js.x[perm], js.v[perm], js.r2[perm])
        a1, j1, n1, l1 = c.query(np.arange(js.n, dtype=np.int32))
    # new particle k is old particle perm[k]
    assert np.array_equal(n1, n0[perm])
    for k in range(0, js.n, 97):
        cnt = n1[k]
        got = l1[k, :cnt]
        assert np.all(np.diff(got) > 0)
        assert np.array_equal(np.sort(perm[got]), l0[perm[k], :cnt])
    err = force_errors(a1, j1, a0[perm], j0[perm])
    say(f"[permuted] a_rel={err['a_rel']} jdot_rel={err['jdot_rel']}")
    # two FP32-term evaluations, each within ~5e-7 of FP64 for a
    assert err["a_rel"]["p99"] <= 1e-6 and err["a_rel"]["max"] <= 2e-6
    assert err["jdot_rel"]["p99"] <= 1e-6


@pytest.mark.parametrize("seed", [1, 2])
def test_king_262144_checked_every_target(ctx, seed):
    """Checked precision on the full King workload (two realisations): every
    target within 1e-6 relative (a and jdot) of the FP64 kernel, over the
    rows whose regular set the FP64 test agrees with (the rest differ by a
    band pair)."""
    say(f"case: king_w6_binaries N=262144 seed {seed} checked, every target")
    js = workloads.king_binaries(262144, seed=seed)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.arange(js.n, dtype=np.int32)
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        chk = ctx.query(t)
        nre = ctx.last_recheck()
        ctx.set_precision(regf.PREC_FP64)
        f64 = ctx.query(t)
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    rows = _same_set_rows(chk[2], chk[3], f64[2], f64[3])
    err = force_errors(chk[0][rows], chk[1][rows], f64[0][rows], f64[1][rows])
    say(f"[king262144 checked vs fp64] recomputed={nre} band_rows={(~rows).sum()} "
        f"a_rel={err['a_rel']} jdot_rel={err['jdot_rel']}")
    assert (~rows).sum() <= 16
    assert err["a_rel"]["max"] <= CHECK_TOL and err["jdot_rel"]["max"] <= CHECK_TOL


@pytest.mark.parametrize("name", ["plummer_n64_seed1.npz", "plummer_n1024_seed1.npz"])
def test_golden_fixtures(name):
    """The HIP path against the committed golden vectors (no oracle call)."""
    import os
    say(f"case: golden {name}")
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))
    cap = int(g["cap"])
    with regf.Context(device=0, cap=cap) as c:
        c.upload(g["m"], g["x"], g["v"], g["r2"])
        acc, adot, nnb, lst = c.query(g["targets"])
    assert np.array_equal(nnb, g["nnb"])
    for k in range(len(nnb)):
        assert np.array_equal(lst[k, :nnb[k]], g["lists"][k, :nnb[k]])
    err = force_errors(acc, adot, g["acc"], g["adot"])
    say(f"[golden {name}] {err}")
    # the default precision: every target within 1e-6
    assert err["a_rel"]["max"] <= 1e-6
    assert err["jdot_rel"]["max"] <= 1e-6


def test_autotune_table_roundtrip(tmp_path):
    """The autotuner installs a 32-bucket table, tuned plans still meet the
    oracle bars, and a saved table loads back."""
    say("case: autotune")
    path = str(tmp_path / "plans.txt")
    try:
        regf.autotune(0, 16384, path)
        lines = [ln for ln in open(path) if not ln.startswith("#")]
        assert len(lines) == 32
        js = workloads.plummer_kroupa(65536)
        with regf.Context(device=0, cap=js.cap) as c:
            c.upload(js.m, js.x, js.v, js.r2)
            for size in (1, 100, 5000):
                t = np.random.default_rng(size).permutation(js.n)[:size].astype(np.int32)
                r = c.query(t)
                say(f"  tuned plan for {size} targets: {c.plan()}")
                check(js, t, r, label=f"tuned/{size}")
        regf.autotune_clear(0)
        regf.autotune_load(0, path)
    finally:
        regf.autotune_clear(0)


def test_hermite_run_one_nbody_time(tmp_path):
    """N=1024 equal-mass Plummer (seed 1, ~48 neighbours) integrated for one
    N-body time with block time steps, the regular force and neighbour lists
    coming from the six legacy entry points: energy is conserved at the same
    level as with the FP64 oracle as provider."""
    from tests import hermite_util
    say("case: Hermite run, N=1024, 1 N-body time")
    js = workloads.plummer_equal(1024)
    ic = str(tmp_path / "ic.bin")
    hermite_util.write_ic(ic, js)
    gpu = hermite_util.run("gpu", ic, 1.0)
    ref = hermite_util.run("oracle", ic, 1.0)
    say(f"  gpu:    {gpu}")
    say(f"  oracle: {ref}")
    assert abs(gpu["dE_over_E"]) < 1e-6
    assert abs(gpu["dE_over_E"]) < 100 * max(abs(ref["dE_over_E"]), 1e-9)
    assert abs(gpu["blocks"] - ref["blocks"]) < 0.05 * ref["blocks"]


def test_hermite_resident_state_matches_host_prediction(tmp_path):
    """The device-resident path (state uploaded once, prediction on the
    device, only the corrected block scattered back) reproduces the legacy
    path, which predicts on the host and uploads all j every block, bit for
    bit over a quarter N-body time."""
    from tests import hermite_util
    say("case: Hermite resident vs legacy")
    js = workloads.plummer_equal(1024)
    ic = str(tmp_path / "ic.bin")
    hermite_util.write_ic(ic, js)
    legacy = hermite_util.run("gpu", ic, 0.25)
    resident = hermite_util.run("resident", ic, 0.25)
    say(f"  legacy:   {legacy}")
    say(f"  resident: {resident}")
    for key in ("E1", "dE_over_E", "blocks", "particle_steps", "com", "x0"):
        assert legacy[key] == resident[key], key


def test_ahmad_cohen_two_time_scales_gpu_vs_oracle(tmp_path):
    # the device-resident path (predict, irregular, regular_step with the
    # neighbour-set correction, derivs23, update_radius, update_state) runs
    # the same Ahmad-Cohen integration as the oracle
    from tests import hermite_util
    js = workloads.plummer_equal(1024)
    ic = str(tmp_path / "ic.bin")
    hermite_util.write_ic(ic, js)
    say("case: Ahmad-Cohen N=1024 t=1 eta 0.02/0.001 gpu")
    g = hermite_util.run_ac("gpu", ic, 1.0, 0.02, 0.001, 48)
    o = hermite_util.run_ac("oracle", ic, 1.0, 0.02, 0.001, 48)
    say(f"  gpu:    {g}")
    say(f"  oracle: {o}")
    assert abs(g["dE_over_E"]) < 1e-6 and g["overflows"] == 0
    assert abs(g["dE_over_E"] - o["dE_over_E"]) < 5e-7
    for k in ("irregular_steps", "regular_steps"):
        assert abs(g[k] - o[k]) <= 0.05 * o[k], (k, g[k], o[k])
