"""GPU parity of the irregular (neighbour) force over resident lists
(abyss_ctx_keep_lists / abyss_ctx_irregular) against the FP64 oracle.

Both sides are FP64; they differ only in summation order and FMA
contraction, so the bar is a normalised error |d a| / sum_j |term_j|
<= 1e-13 per target.  The split is also checked as a whole: the regular
part of a step plus the irregular part over that step's lists equals the
unsplit direct sum to the regular kernel's own accuracy (1e-6).
"""
import sys

import numpy as np
import pytest

import oracle
from abyss_amd import regf, workloads

pytestmark = pytest.mark.gpu

TOL_IRR_NORM = 1e-13


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scales(x, v, m, t, nnb, lst, cap):
    """Per target: sum_j m/r^2 and sum_j 4 m |v|/r^3 over the list."""
    sa = np.zeros(len(t))
    sj = np.zeros(len(t))
    for q, i in enumerate(t):
        c = min(abs(int(nnb[q])), cap)
        j = lst[q, :c]
        d = x[j] - x[i]
        r2 = np.sum(d * d, axis=1)
        ok = r2 > 0
        r2 = np.where(ok, r2, 1.0)
        w = np.where(ok, m[j] / r2, 0.0)
        sa[q] = np.sum(w)
        sj[q] = np.sum(4 * w * np.linalg.norm(v[j] - v[i], axis=1) / np.sqrt(r2))
    return sa, sj


def _norm_err(got, want, scale):
    d = np.linalg.norm(got - want, axis=1)
    return d / np.where(scale > 0, scale, 1.0)


def check_irregular(x, v, m, t, nnb, lst, cap, acc, adot, label):
    ra, rj = oracle.irregular(x, v, m, t, nnb, lst, cap)
    sa, sj = _scales(x, v, m, t, nnb, lst, cap)
    ea = _norm_err(acc, ra, sa)
    ej = _norm_err(adot, rj, sj)
    say(f"[{label}] nt={len(t)} a_norm_max={ea.max():.3e} jdot_norm_max={ej.max():.3e} "
        f"mean_nnb={np.mean(np.abs(nnb)):.1f}")
    assert ea.max() <= TOL_IRR_NORM
    assert ej.max() <= TOL_IRR_NORM
    # a target without neighbours gets exactly zero
    empty = np.asarray(nnb) == 0
    assert np.all(acc[empty] == 0) and np.all(adot[empty] == 0)


@pytest.fixture(scope="module")
def ctx():
    c = regf.Context(device=0, cap=256)
    c.keep_lists(True)
    yield c
    c.close()


@pytest.mark.parametrize("size", [1, 100, 8192, 65536])
def test_irregular_matches_oracle_kroupa_65536(ctx, size):
    say(f"case: irregular plummer_kroupa N=65536 block={size}")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    t = np.random.default_rng(size + 7).permutation(js.n)[:size].astype(np.int32)
    acc_r, adot_r, nnb, lst = ctx.query(t)
    acc_i, adot_i = ctx.irregular(t)
    check_irregular(js.x, js.v, js.m, t, nnb, lst, js.cap, acc_i, adot_i, f"irr/{size}")
    # regular + irregular = unsplit direct sum (oracle with r_nb = 0)
    fa, fj, _, _, sa, sj = oracle.regf(js.x, js.v, js.m, np.zeros(js.n), t, js.cap, scales=True)
    ea = _norm_err(acc_r + acc_i, fa, sa)
    ej = _norm_err(adot_r + adot_i, fj, sj)
    say(f"  reg+irr vs unsplit: a_norm_max={ea.max():.3e} jdot_norm_max={ej.max():.3e}")
    assert ea.max() <= 1e-6 and ej.max() <= 1e-6
    # bitwise reproducible
    a2, j2 = ctx.irregular(t)
    assert np.array_equal(a2, acc_i) and np.array_equal(j2, adot_i)


def test_irregular_uses_each_particles_last_list(ctx):
    say("case: irregular last list per particle, plummer_equal N=1024")
    js = workloads.plummer_equal(1024)
    ctx.upload(js.m, js.x, js.v, js.r2)
    rng = np.random.default_rng(3)
    perm = rng.permutation(js.n).astype(np.int32)
    t1, t2 = perm[:600], perm[400:900]   # 200 particles in both launches
    _, _, n1, l1 = ctx.query(t1)
    # second launch with wider radii: the overlap's lists change
    r2b = js.r2 * 1.7
    ctx.upload(js.m, js.x, js.v, r2b)
    _, _, n2, l2 = ctx.query(t2)
    # the last list of each particle; never-targeted particles have none
    cap = js.cap
    last_n = np.zeros(js.n, dtype=np.int32)
    last_l = np.zeros((js.n, cap), dtype=np.int32)
    last_n[t1], last_l[t1] = n1, l1
    last_n[t2], last_l[t2] = n2, l2
    q = np.arange(js.n, dtype=np.int32)
    acc, adot = ctx.irregular(q)
    check_irregular(js.x, js.v, js.m, q, last_n[q], last_l[q], cap, acc, adot, "irr/last")
    assert np.sum(last_n == 0) >= 100   # untargeted: exactly zero (checked above)


def test_irregular_on_predicted_positions(ctx):
    say("case: irregular after device prediction, plummer_equal N=1024")
    js = workloads.plummer_equal(1024)
    rng = np.random.default_rng(11)
    a = rng.normal(size=(js.n, 3)) * 0.1
    ad = rng.normal(size=(js.n, 3)) * 0.1
    t0 = rng.uniform(0.0, 0.01, js.n)
    ctx.upload_state(js.m, js.x, js.v, a, ad, t0, js.r2)
    ctx.predict(0.01)
    t = rng.permutation(js.n)[:512].astype(np.int32)
    _, _, nnb, lst = ctx.query(t)
    ctx.predict(0.0125)   # irregular step later than the regular one
    acc, adot = ctx.irregular(t)
    xp, vp = oracle.predict2(js.x, js.v, a, ad, 0.0125 - t0)
    check_irregular(xp, vp, js.m, t, nnb, lst, js.cap, acc, adot, "irr/predicted")


def test_irregular_overflowed_lists_are_truncated():
    say("case: irregular with overflowed lists, cap=16")
    js = workloads.plummer_equal(1024)
    with regf.Context(device=0, cap=16) as small:
        small.keep_lists(True)
        small.upload(js.m, js.x, js.v, js.r2)
        t = np.arange(0, js.n, 5, dtype=np.int32)
        _, _, nnb, lst = small.query(t)
        assert small.last_status == regf.ERR_NEIGHBOR_OVERFLOW and np.all(nnb < 0)
        acc, adot = small.irregular(t)
        check_irregular(js.x, js.v, js.m, t, nnb, lst, 16, acc, adot, "irr/overflow")


def test_irregular_errors_and_resets():
    say("case: irregular errors and resets")
    js = workloads.plummer_equal(1024)
    with regf.Context(device=0, cap=256) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        with pytest.raises(regf.AbyssError) as e:
            c.irregular([0])
        assert e.value.code == regf.ERR_NO_JSET
        c.keep_lists(True)
        acc, adot = c.irregular([0, 5])          # never targeted: empty lists
        assert np.all(acc == 0) and np.all(adot == 0)
        c.query(np.arange(64, dtype=np.int32))
        with pytest.raises(regf.AbyssError) as e:
            c.irregular([js.n])
        assert e.value.code == regf.ERR_TARGET_RANGE
        with pytest.raises(regf.AbyssError) as e:
            c.irregular([])
        assert e.value.code == regf.ERR_INVALID_ARG
        acc, _ = c.irregular([3])
        assert np.linalg.norm(acc) > 0
        # a j-set of another size empties every list
        c.upload(js.m[:512], js.x[:512], js.v[:512], js.r2[:512])
        acc, adot = c.irregular([3])
        assert np.all(acc == 0) and np.all(adot == 0)
        c.keep_lists(False)
        with pytest.raises(regf.AbyssError):
            c.irregular([3])


def _step_case(ctx, js, t, r2_new, dt, a_old, label):
    """Regular step after a first launch with js.r2 (the old lists): the GPU
    correction against the oracle's, fed with the GPU's own lists and new
    regular force (so only the correction itself is compared)."""
    _, _, no, lo = ctx.query(t)                 # old lists, kept resident
    ctx.upload(js.m, js.x, js.v, r2_new)        # same n: resident lists stay
    acc, adot, nn, ln, out = ctx.regular_step(t, dt, a_old)
    ref = oracle.regular_correct(js.x, js.v, js.m, t, dt, a_old, acc, adot, no, lo, nn, ln,
                                 js.cap)
    # summation order differs (lanes + butterfly vs list order): bound by the
    # summed magnitudes of every term either list contributes
    so, _ = _scales(js.x, js.v, js.m, t, no, lo, js.cap)
    sn, sjn = _scales(js.x, js.v, js.m, t, nn, ln, js.cap)
    scale = (np.linalg.norm(a_old[:, :3], axis=1) + np.linalg.norm(acc, axis=1) + so + sn)
    e2 = np.linalg.norm(out[:, 6:9] - ref[:, 6:9], axis=1) * dt ** 2 / scale
    e3 = np.linalg.norm(out[:, 9:12] - ref[:, 9:12], axis=1) * dt ** 3 / scale
    ex = np.linalg.norm(out[:, 0:3] - ref[:, 0:3], axis=1)
    ev = np.linalg.norm(out[:, 3:6] - ref[:, 3:6], axis=1)
    ei = _norm_err(out[:, 12:15], ref[:, 12:15], sn)
    ej = _norm_err(out[:, 15:18], ref[:, 15:18], sjn)
    gained = sum(len(set(ln[q, :abs(nn[q])]) - set(lo[q, :abs(no[q])])) for q in range(len(t)))
    lost = sum(len(set(lo[q, :abs(no[q])]) - set(ln[q, :abs(nn[q])])) for q in range(len(t)))
    say(f"[{label}] nt={len(t)} gained={gained} lost={lost} a2_err={e2.max():.2e} "
        f"a3_err={e3.max():.2e} x_err={ex.max():.2e} v_err={ev.max():.2e} "
        f"irr={ei.max():.2e}/{ej.max():.2e}")
    assert e2.max() <= 1e-13 and e3.max() <= 1e-12
    xs = np.abs(js.x[t]).max() + 1.0
    vs = np.abs(js.v[t]).max() + 1.0
    assert ex.max() <= 1e-14 * xs and ev.max() <= 1e-14 * vs
    assert ei.max() <= TOL_IRR_NORM and ej.max() <= TOL_IRR_NORM
    # the resident lists are now the new ones: the irregular force over them
    # is the step's a_irr, bit for bit (same terms, same lane order)
    ai, ji = ctx.irregular(t)
    assert np.array_equal(ai, out[:, 12:15]) and np.array_equal(ji, out[:, 15:18])
    return gained, lost


@pytest.mark.parametrize("size", [1, 100, 8192])
def test_regular_step_matches_oracle_kroupa_65536(ctx, size):
    say(f"case: regular_step plummer_kroupa N=65536 block={size}")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    rng = np.random.default_rng(size + 31)
    t = rng.permutation(js.n)[:size].astype(np.int32)
    r2_new = js.r2 * np.where(rng.uniform(size=js.n) < 0.5, 1.4, 0.7)
    dt = rng.uniform(1e-3, 1e-2, size)
    a_old = rng.normal(size=(size, 6))
    gained, lost = _step_case(ctx, js, t, r2_new, dt, a_old, f"step/{size}")
    if size >= 100:
        assert gained > 0 and lost > 0


def test_regular_step_checked_precision(ctx):
    """The fused step under checked precision: its regular force is the
    checked query's, bit for bit, and the correction matches the oracle."""
    say("case: regular_step checked precision plummer_kroupa N=65536 block=4096")
    js = workloads.plummer_kroupa(65536)
    ctx.upload(js.m, js.x, js.v, js.r2)
    rng = np.random.default_rng(4127)
    t = rng.permutation(js.n)[:4096].astype(np.int32)
    r2_new = js.r2 * np.where(rng.uniform(size=js.n) < 0.5, 1.4, 0.7)
    dt = rng.uniform(1e-3, 1e-2, t.size)
    a_old = rng.normal(size=(t.size, 6))
    ctx.set_precision(regf.PREC_CHECKED)
    try:
        _step_case(ctx, js, t, r2_new, dt, a_old, "step/checked")
        acc, adot, nn, _, _ = ctx.regular_step(t, dt, a_old)
        n_step = ctx.last_recheck()
        qa, qj, qn, _ = ctx.query(t)
        assert ctx.last_recheck() == n_step > 0
    finally:
        ctx.set_precision(regf.PREC_SPLIT)
    assert np.array_equal(acc, qa) and np.array_equal(adot, qj) and np.array_equal(nn, qn)


def test_regular_step_first_step_and_overflow():
    say("case: regular_step without old lists, and with overflowed lists (cap=16)")
    js = workloads.plummer_equal(1024)
    rng = np.random.default_rng(5)
    t = np.arange(0, js.n, 3, dtype=np.int32)
    dt = np.full(t.size, 0.01)
    a_old = rng.normal(size=(t.size, 6))
    with regf.Context(device=0, cap=256) as c:
        c.keep_lists(True)
        c.upload(js.m, js.x, js.v, js.r2)
        # no old lists: every new neighbour is gained
        acc, adot, nn, ln, out = c.regular_step(t, dt, a_old)
        empty = np.zeros((t.size, 256), dtype=np.int32)
        ref = oracle.regular_correct(js.x, js.v, js.m, t, dt, a_old, acc, adot,
                                     np.zeros(t.size, np.int32), empty, nn, ln, 256)
        np.testing.assert_allclose(out[:, 12:], ref[:, 12:], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(out[:, 6:12], ref[:, 6:12], rtol=1e-9, atol=1e-9)
    with regf.Context(device=0, cap=16) as c:
        c.keep_lists(True)
        c.upload(js.m, js.x, js.v, js.r2)
        _, _, no, lo = c.query(t)
        c.upload(js.m, js.x, js.v, js.r2 * 1.3)
        acc, adot, nn, ln, out = c.regular_step(t, dt, a_old)
        assert c.last_status == regf.ERR_NEIGHBOR_OVERFLOW and np.all(nn < 0)
        ref = oracle.regular_correct(js.x, js.v, js.m, t, dt, a_old, acc, adot, no, lo, nn, ln,
                                     16)
        np.testing.assert_allclose(out[:, 12:], ref[:, 12:], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(out[:, 6:12], ref[:, 6:12], rtol=1e-9, atol=1e-9)


def test_regular_step_errors():
    say("case: regular_step invalid inputs")
    js = workloads.plummer_equal(1024)
    with regf.Context(device=0, cap=256) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        t = np.arange(8, dtype=np.int32)
        a_old = np.zeros((8, 6))
        with pytest.raises(regf.AbyssError) as e:
            c.regular_step(t, 0.01, a_old)
        assert e.value.code == regf.ERR_NO_JSET
        c.keep_lists(True)
        for bad_dt in (0.0, -1.0, np.nan):
            with pytest.raises(regf.AbyssError) as e:
                c.regular_step(t, bad_dt, a_old)
            assert e.value.code == regf.ERR_INVALID_INPUT
        bad = a_old.copy()
        bad[3, 2] = np.inf
        with pytest.raises(regf.AbyssError) as e:
            c.regular_step(t, 0.01, bad)
        assert e.value.code == regf.ERR_INVALID_INPUT
        with pytest.raises(regf.AbyssError) as e:
            c.regular_step(np.array([0, js.n], dtype=np.int32), 0.01, np.zeros((2, 6)))
        assert e.value.code == regf.ERR_TARGET_RANGE
        # still usable afterwards
        c.regular_step(t, 0.01, a_old)
