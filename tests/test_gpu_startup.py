"""GPU parity of the start-up derivatives (abyss_ctx_derivs23) and the
neighbour-radius rule (abyss_ctx_update_radius) against the oracle."""
import sys

import numpy as np
import pytest

import oracle
from abyss_amd import regf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(n, seed):
    js = workloads.plummer_equal(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    # a, adot: the true total acceleration and jerk (oracle with r_nb = 0)
    # plus a perturbation, so relative accelerations are not tiny
    t = np.arange(js.n, dtype=np.int32)
    a, ad, _, _ = oracle.regf(js.x, js.v, js.m, np.zeros(js.n), t, 4)
    a = a + rng.normal(size=a.shape) * 0.1
    ad = ad + rng.normal(size=ad.shape) * 0.1
    return js, a, ad


@pytest.mark.parametrize("ref", [True, False])
@pytest.mark.parametrize("size", [1, 100, 2048])
def test_derivs23_matches_oracle(ref, size):
    say(f"case: derivs23 plummer N=2048 block={size} reference_signs={ref}")
    js, a, ad = _state(2048, 3)
    t = np.random.default_rng(size).permutation(js.n)[:size].astype(np.int32)
    with regf.Context(device=0, cap=256) as c:
        c.upload_state(js.m, js.x, js.v, a, ad, np.zeros(js.n), js.r2)
        got = c.derivs23(t, reference_signs=ref)
        again = c.derivs23(t, reference_signs=ref)
    want = oracle.derivs23(js.x, js.v, a, ad, js.m, js.r2, t, ref_signs=ref)
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(got - want) / scale
    say(f"  max err / row max = {err.max():.3e}; irregular part nonzero: "
        f"{np.abs(want[:, 6:]).max() > 0}")
    assert err.max() <= 1e-10
    assert np.array_equal(got, again)


def test_derivs23_signs_differ_and_errors():
    js, a, ad = _state(1024, 5)
    t = np.arange(0, js.n, 7, dtype=np.int32)
    with regf.Context(device=0, cap=256) as c:
        with pytest.raises(regf.AbyssError) as e:
            c.derivs23(t)
        assert e.value.code == regf.ERR_NO_JSET
        c.upload_state(js.m, js.x, js.v, a, ad, np.zeros(js.n), js.r2)
        r = c.derivs23(t, True)
        x = c.derivs23(t, False)
        assert np.abs(r - x).max() > 1e-6 * np.abs(x).max()
        with pytest.raises(regf.AbyssError) as e:
            c.derivs23([js.n])
        assert e.value.code == regf.ERR_TARGET_RANGE


def test_update_radius_matches_host_rule_and_feeds_next_launch():
    say("case: update_radius plummer N=4096")
    js = workloads.plummer_equal(4096, k_nb=48)
    rng = np.random.default_rng(9)
    # radii spread so counts land above, below and at fix
    r2 = js.r2 * rng.uniform(0.3, 3.0, js.n)
    t = rng.permutation(js.n)[:1500].astype(np.int32)
    fix, nmax = 48, 200
    with regf.Context(device=0, cap=256) as c:
        c.upload(js.m, js.x, js.v, r2)
        _, _, nnb, _ = c.query(t)
        got = c.update_radius(fix, nmax)
        want = oracle.update_radius(np.abs(nnb), r2[t], fix, nmax)
        assert np.array_equal(got, want)
        assert (np.abs(nnb) > fix).any() and (np.abs(nnb) < fix).any()
        # the next launch uses the new radii
        r2_new = r2.copy()
        r2_new[t] = got
        _, _, nnb2, lst2 = c.query(t)
        _, _, rn, rl = oracle.regf(js.x, js.v, js.m, r2_new, t, 256)
        from oracle.compare import compare_lists
        res = compare_lists(nnb2, lst2, rn, rl, js.x, r2_new, t, band=1e-6)
        say(f"  lists after update: {res}")
        assert res["hard_mismatches"] == 0


def test_update_radius_needs_a_launch_of_the_current_targets():
    js = workloads.plummer_equal(1024)
    with regf.Context(device=0, cap=256) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        c.set_targets(np.arange(10, dtype=np.int32))
        with pytest.raises(regf.AbyssError) as e:
            c.update_radius(48, 256)
        assert e.value.code == regf.ERR_NO_JSET
        c.launch()
        c.sync()
        assert c.update_radius(48, 256).shape == (10,)
        with pytest.raises(regf.AbyssError) as e:
            c.update_radius(48, 48)          # nmax must exceed fix
        assert e.value.code == regf.ERR_INVALID_ARG
