"""Pins the FP64 oracle against closed forms, calculus identities, invariants
and an independent brute force.  Each check is chosen so that a plausible
mistake in the restated formula (sign, power of r, the factor 3 in the jerk,
the <= boundary, self-exclusion, list order) fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle


def _rand_cluster(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)) * spread
    v = rng.normal(size=(n, 3)) * 0.3
    m = rng.uniform(0.5, 2.0, n) / n
    return x, v, m


def test_two_body_closed_form():
    # r = x1 - x0 = (3,4,0), |r| = 5; v = v1 - v0 = (0.3,0.3,-0.8), r.v = 2.1
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    v = np.array([[0.1, -0.2, 0.3], [0.4, 0.1, -0.5]])
    m = np.array([1.5, 2.5])
    acc, adot, nnb, _ = oracle.regf(x, v, m, np.zeros(2), [0, 1], cap=4)
    # hand-evaluated: a0 = 2.5 (3,4,0)/125 ; jdot0 = 2.5[(v)/125 - 3*2.1 (3,4,0)/3125]
    np.testing.assert_allclose(acc[0], [0.06, 0.08, 0.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(adot[0], [-0.00912, -0.01416, -0.016], rtol=0, atol=1e-15)
    np.testing.assert_allclose(acc[1], [-0.036, -0.048, 0.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(adot[1], [0.005472, 0.008496, 0.0096], rtol=0, atol=1e-15)
    assert list(nnb) == [0, 0]


def _potential(x, m, i, xi):
    d = x - xi
    r = np.sqrt(np.sum(d * d, axis=1))
    r[i] = np.inf
    return -np.sum(m / r)


def test_acc_is_minus_grad_potential():
    x, v, m = _rand_cluster(32, 3)
    acc, _, _, _ = oracle.regf(x, v, m, np.zeros(32), np.arange(32), cap=32)
    h = 1e-4
    for i in range(0, 32, 5):
        g = np.zeros(3)
        for d in range(3):
            e = np.zeros(3)
            e[d] = 1.0
            # 4th-order central difference
            f = [_potential(x, m, i, x[i] + s * h * e) for s in (-2, -1, 1, 2)]
            g[d] = (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h)
        np.testing.assert_allclose(acc[i], -g, rtol=1e-7, atol=1e-9 * np.abs(acc[i]).max())


def test_jerk_is_time_derivative_of_acc():
    # Straight-line motion x_j(t) = x_j + v_j t: jdot_i = d a_i / dt at t = 0.
    x, v, m = _rand_cluster(24, 5)
    n = len(m)
    t = np.arange(n)
    _, adot, _, _ = oracle.regf(x, v, m, np.zeros(n), t, cap=n)
    h = 1e-4
    a = {}
    for s in (-2, -1, 1, 2):
        a[s] = oracle.regf(x + s * h * v, v, m, np.zeros(n), t, cap=n)[0]
    fd = (a[-2] - 8 * a[-1] + 8 * a[1] - a[2]) / (12 * h)
    np.testing.assert_allclose(adot, fd, rtol=1e-6, atol=1e-8 * np.abs(adot).max())


def test_momentum_conservation_full_force():
    x, v, m = _rand_cluster(200, 7)
    n = len(m)
    acc, adot, _, _ = oracle.regf(x, v, m, np.zeros(n), np.arange(n), cap=n)
    pa = (m[:, None] * acc).sum(0)
    pj = (m[:, None] * adot).sum(0)
    assert np.all(np.abs(pa) < 1e-13 * np.sum(m[:, None] * np.abs(acc)))
    assert np.all(np.abs(pj) < 1e-13 * np.sum(m[:, None] * np.abs(adot)))


def test_homogeneity_scalings():
    x, v, m = _rand_cluster(64, 11)
    n = len(m)
    r2 = np.full(n, 0.3)
    t = np.arange(n)
    a0, j0, c0, l0 = oracle.regf(x, v, m, r2, t, cap=n)
    lam = 3.0
    a1, j1, c1, l1 = oracle.regf(x, v, lam * m, r2, t, cap=n)
    np.testing.assert_allclose(a1, lam * a0, rtol=1e-13)
    np.testing.assert_allclose(j1, lam * j0, rtol=1e-13)
    assert np.array_equal(c0, c1)
    # x -> lam x (and r_nb^2 -> lam^2 r_nb^2): a -> a/lam^2, jdot -> jdot/lam^3
    a2, j2, c2, l2 = oracle.regf(lam * x, v, m, lam * lam * r2, t, cap=n)
    np.testing.assert_allclose(a2, a0 / lam ** 2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(j2, j0 / lam ** 3, rtol=1e-12, atol=1e-15)
    assert np.array_equal(c0, c2) and np.array_equal(l0, l2)
    # v -> lam v: a unchanged, jdot -> lam jdot
    a3, j3, _, _ = oracle.regf(x, lam * v, m, r2, t, cap=n)
    np.testing.assert_allclose(a3, a0, rtol=0, atol=0)
    np.testing.assert_allclose(j3, lam * j0, rtol=1e-13)


def test_radius_limits():
    x, v, m = _rand_cluster(50, 13)
    n = len(m)
    t = np.arange(n)
    a0, j0, c0, _ = oracle.regf(x, v, m, np.zeros(n), t, cap=n)
    assert np.all(c0 == 0)
    ainf, jinf, cinf, linf = oracle.regf(x, v, m, np.full(n, np.inf), t, cap=n)
    assert np.all(ainf == 0) and np.all(jinf == 0)
    assert np.all(cinf == n - 1)
    for k in range(n):
        assert list(linf[k, :n - 1]) == [j for j in range(n) if j != k]
    # a_reg(r_nb) + sum over the listed neighbours of the pair force = full force
    r2 = np.full(n, 0.8)
    ar, jr, cr, lr = oracle.regf(x, v, m, r2, t, cap=n)
    for k in range(n):
        nb = lr[k, :cr[k]]
        d = x[nb] - x[k]
        r = np.sqrt(np.sum(d * d, axis=1))
        pair = (m[nb] / r ** 3)[:, None] * d
        np.testing.assert_allclose(ar[k] + pair.sum(0), a0[k], rtol=1e-12, atol=1e-14)


def test_boundary_is_inclusive_and_self_excluded_by_index():
    # r^2 = 0.75^2 + 1^2 = 1.5625 exactly representable
    x = np.array([[0.0, 0.0, 0.0], [0.75, 1.0, 0.0], [5.0, 0.0, 0.0]])
    v = np.zeros((3, 3))
    m = np.array([1.0, 1.0, 1.0])
    r2 = np.array([1.5625, 0.0, 0.0])
    _, _, c, lst = oracle.regf(x, v, m, r2, [0], cap=3)
    assert c[0] == 1 and lst[0, 0] == 1          # r^2 == r_nb^2 -> neighbour
    r2[0] = np.nextafter(1.5625, 0.0)
    a, _, c, _ = oracle.regf(x, v, m, r2, [0], cap=3)
    assert c[0] == 0 and a[0, 0] > 0            # r^2 > r_nb^2 -> regular


def test_coincident_particle_is_neighbour_without_nan():
    x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [4.0, 2.0, 3.0]])
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    m = np.array([1.0, 2.0, 3.0])
    a, j, c, lst = oracle.regf(x, v, m, np.zeros(3), [0, 1], cap=3)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(j))
    assert list(c) == [1, 1] and lst[0, 0] == 1 and lst[1, 0] == 0
    # only particle 2 (at distance 3 along x) contributes
    np.testing.assert_allclose(a[0], [3.0 / 9.0, 0, 0], rtol=1e-15)


def test_overflow_convention():
    x, v, m = _rand_cluster(40, 17)
    n = len(m)
    _, _, cfull, lfull = oracle.regf(x, v, m, np.full(n, np.inf), [3], cap=n)
    _, _, c, lst = oracle.regf(x, v, m, np.full(n, np.inf), [3], cap=5)
    assert cfull[0] == n - 1 and c[0] == -(n - 1)
    assert list(lst[0]) == list(lfull[0, :5])


def test_matches_independent_numpy_brute_force():
    x, v, m = _rand_cluster(150, 19)
    n = len(m)
    rng = np.random.default_rng(23)
    r2 = rng.uniform(0.05, 2.0, n)
    t = rng.permutation(n)[:97]
    acc, adot, c, lst = oracle.regf(x, v, m, r2, t, cap=n)
    D = x[None, :, :] - x[t][:, None, :]                # [nt, n, 3]
    V = v[None, :, :] - v[t][:, None, :]
    R2 = np.sum(D * D, axis=2)
    far = R2 > r2[t][:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(far, 1.0 / np.sqrt(R2), 0.0)
    w = m[None, :] * inv ** 3
    RV = np.sum(D * V, axis=2)
    a_np = np.sum(w[:, :, None] * D, axis=1)
    j_np = np.sum(w[:, :, None] * (V - 3.0 * (RV * inv * inv)[:, :, None] * D), axis=1)
    np.testing.assert_allclose(acc, a_np, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(adot, j_np, rtol=1e-11, atol=1e-13)
    for q, i in enumerate(t):
        want = [j for j in range(n) if j != i and not far[q, j]]
        assert c[q] == len(want)
        assert list(lst[q, :c[q]]) == want


def test_plummer_radial_acceleration_matches_analytic_profile():
    from abyss_amd import ic
    x, v, m = ic.plummer(4096, seed=2)
    n = len(m)
    t = np.arange(0, n, 4)
    acc, _, _, _ = oracle.regf(x, v, m, np.zeros(n), t, cap=1, nthreads=0)
    a_pl = 3 * math.pi / 16
    r = np.linalg.norm(x[t], axis=1)
    ar = np.sum(acc * x[t], axis=1) / r          # radial component
    model = -r / (r * r + a_pl * a_pl) ** 1.5    # G = M = 1
    sel = (r > 0.3) & (r < 2.0)
    ratio = ar[sel] / model[sel]
    assert abs(np.median(ratio) - 1.0) < 0.05


def test_band_pairs():
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    r2 = np.array([1.0 * (1 + 5e-7), 0.0, 0.0])
    bc, bl = oracle.band_pairs(x, r2, [0], rel=1e-6)
    assert bc[0] == 2 and list(bl[0, :2]) == [1, 2]
    r2[0] = 1.0 * (1 + 3e-6)
    bc, _ = oracle.band_pairs(x, r2, [0], rel=1e-6)
    assert bc[0] == 0


def test_predictor_matches_taylor_series():
    xp, vp = oracle.predict2([[1.0, 0, 0]], [[2.0, 0, 0]], [[3.0, 0, 0]], [[4.0, 0, 0]], 0.5)
    # x + v dt + a dt^2/2 + adot dt^3/6 ; v + a dt + adot dt^2/2
    assert xp[0, 0] == pytest.approx(1 + 1 + 0.375 + 4 * 0.125 / 6, rel=1e-15)
    assert vp[0, 0] == pytest.approx(2 + 1.5 + 0.5, rel=1e-15)
    xp, vp = oracle.predict2([[1.0, 2, 3]], [[2.0, 0, 0]], [[3.0, 0, 0]], [[4.0, 0, 0]], 0.0)
    assert list(xp[0]) == [1.0, 2, 3] and list(vp[0]) == [2.0, 0, 0]


def test_irregular_two_body_closed_form():
    # same pair as the regular closed form, but j is a neighbour: the force
    # moves from the regular to the irregular part unchanged
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    v = np.array([[0.1, -0.2, 0.3], [0.4, 0.1, -0.5]])
    m = np.array([1.5, 2.5])
    _, _, nnb, lst = oracle.regf(x, v, m, np.full(2, 26.0), [0, 1], cap=4)
    assert list(nnb) == [1, 1]
    acc, adot = oracle.irregular(x, v, m, [0, 1], nnb, lst, 4)
    np.testing.assert_allclose(acc[0], [0.06, 0.08, 0.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(adot[0], [-0.00912, -0.01416, -0.016], rtol=0, atol=1e-15)
    np.testing.assert_allclose(acc[1], [-0.036, -0.048, 0.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(adot[1], [0.005472, 0.008496, 0.0096], rtol=0, atol=1e-15)


def test_regular_plus_irregular_is_unsplit_brute_force():
    # a_reg + a_irr over the list of the same step = the direct sum over all
    # j != i, computed independently in numpy
    x, v, m = _rand_cluster(200, 41)
    n = len(m)
    rng = np.random.default_rng(43)
    r2 = rng.uniform(0.05, 3.0, n)
    t = rng.permutation(n)[:120]
    acc, adot, c, lst = oracle.regf(x, v, m, r2, t, cap=n)
    assert c.max() > 0 and c.max() < n - 1 and np.mean(c > 0) > 0.5
    ai, ji = oracle.irregular(x, v, m, t, c, lst, n)
    D = x[None, :, :] - x[t][:, None, :]
    V = v[None, :, :] - v[t][:, None, :]
    R2 = np.sum(D * D, axis=2)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(R2 > 0, 1.0 / np.sqrt(R2), 0.0)
    w = m[None, :] * inv ** 3
    RV = np.sum(D * V, axis=2)
    a_np = np.sum(w[:, :, None] * D, axis=1)
    j_np = np.sum(w[:, :, None] * (V - 3.0 * (RV * inv * inv)[:, :, None] * D), axis=1)
    np.testing.assert_allclose(acc + ai, a_np, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(adot + ji, j_np, rtol=1e-11, atol=1e-13)
    # and the irregular part alone is the brute force restricted to the list
    near = np.zeros_like(R2, dtype=bool)
    for q in range(len(t)):
        near[q, lst[q, :c[q]]] = True
    a_n = np.sum((w * near)[:, :, None] * D, axis=1)
    np.testing.assert_allclose(ai, a_n, rtol=1e-12, atol=1e-14)


def test_irregular_empty_truncated_and_coincident():
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 0, 0], [0.0, 2.0, 0]])
    v = np.zeros((4, 3))
    m = np.array([1.0, 2.0, 3.0, 4.0])
    # target 0: coincident j=2 adds nothing; -3 with cap 2 sums the first 2
    lists = np.array([[1, 2], [0, 0]], dtype=np.int32)
    acc, _ = oracle.irregular(x, v, m, [0, 1], [-3, 0], lists, 2)
    np.testing.assert_allclose(acc[0], [2.0, 0, 0], rtol=0, atol=0)
    np.testing.assert_array_equal(acc[1], [0.0, 0, 0])
    lists = np.array([[1, 2, 3]], dtype=np.int32)
    acc, _ = oracle.irregular(x, v, m, [0], [3], lists, 3)
    np.testing.assert_allclose(acc[0], [2.0, 1.0, 0], rtol=0, atol=1e-15)


def test_regular_corrector_cubic_closed_form():
    # no neighbours (a_tmp = 0): for a regular force that is a cubic in time,
    # a(t) = A0 + A1 t + A2 t^2/2 + A3 t^3/6, the corrector recovers a2 = A2,
    # a3 = A3 and moves the predicted x, v by the remaining Taylor terms
    x = np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])
    v = np.array([[0.1, 0.2, 0.3], [0.0, 0.0, 0.0]])
    m = np.array([1.0, 1.0])
    A0, A1 = np.array([0.3, -0.2, 0.1]), np.array([1.5, 0.25, -0.75])
    A2, A3 = np.array([-2.0, 0.5, 4.0]), np.array([8.0, -3.0, 1.0])
    dt = 0.125
    a1 = A0 + A1 * dt + A2 * dt ** 2 / 2 + A3 * dt ** 3 / 6
    j1 = A1 + A2 * dt + A3 * dt ** 2 / 2
    empty = np.zeros((1, 4), dtype=np.int32)
    out = oracle.regular_correct(x, v, m, [0], dt, np.concatenate([A0, A1])[None], a1[None],
                                 j1[None], [0], empty, [0], empty, 4)[0]
    np.testing.assert_allclose(out[6:9], A2, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[9:12], A3, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(out[0:3] - x[0], A2 * dt ** 4 / 24 + A3 * dt ** 5 / 120,
                               rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(out[3:6] - v[0], A2 * dt ** 3 / 6 + A3 * dt ** 4 / 24,
                               rtol=1e-10, atol=1e-16)
    assert np.all(out[12:] == 0)


def test_regular_corrector_is_invariant_under_the_split():
    # Moving pairs between the regular and irregular parts must not change
    # the corrector: the step (L_old -> L_new, new force of the L_new split)
    # equals the step (L_old -> L_old, new force of the L_old split), because
    # a_tmp converts the new force back to the old neighbour set.  Radii grow
    # for half the targets and shrink for the rest, so pairs are both gained
    # and lost.
    x, v, m = _rand_cluster(300, 61)
    n = len(m)
    rng = np.random.default_rng(67)
    r2_old = rng.uniform(0.1, 1.5, n)
    r2_new = r2_old * np.where(rng.uniform(size=n) < 0.5, 1.6, 0.55)
    t = rng.permutation(n)[:150]
    ao, jo, no, lo = oracle.regf(x, v, m, r2_old, t, cap=n)
    an, jn, nn, ln = oracle.regf(x, v, m, r2_new, t, cap=n)
    gained = sum(len(set(ln[q, :nn[q]]) - set(lo[q, :no[q]])) for q in range(len(t)))
    lost = sum(len(set(lo[q, :no[q]]) - set(ln[q, :nn[q]])) for q in range(len(t)))
    assert gained > 100 and lost > 100
    dt = rng.uniform(0.01, 0.05, len(t))
    a_old = np.concatenate([ao, jo], axis=1) + rng.normal(size=(len(t), 6)) * 0.01
    o1 = oracle.regular_correct(x, v, m, t, dt, a_old, an, jn, no, lo, nn, ln, n)
    o2 = oracle.regular_correct(x, v, m, t, dt, a_old, ao, jo, no, lo, no, lo, n)
    sa = np.abs(ao).max()
    np.testing.assert_allclose(o1[:, 6:9], o2[:, 6:9], rtol=0, atol=1e-11 * sa / dt.min() ** 2)
    np.testing.assert_allclose(o1[:, 9:12], o2[:, 9:12], rtol=0, atol=1e-11 * sa / dt.min() ** 3)
    np.testing.assert_allclose(o1[:, 0:6], o2[:, 0:6], rtol=1e-13, atol=1e-15)
    # the irregular part is the sum over the new list
    ai, ji = oracle.irregular(x, v, m, t, nn, ln, n)
    np.testing.assert_allclose(o1[:, 12:15], ai, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o1[:, 15:18], ji, rtol=1e-11, atol=1e-13)
    # and with the old split the correction vanishes: a2 from a0 - a1 alone
    da = (a_old[:, :3] - ao) / dt[:, None] ** 2
    ad = (a_old[:, 3:] + jo) / dt[:, None]
    np.testing.assert_allclose(o2[:, 6:9], -6 * da - 2 * ad - 2 * a_old[:, 3:] / dt[:, None],
                               rtol=1e-12, atol=1e-12 * np.abs(da).max())


def _pair_force(m, x):
    r2 = np.sum(x * x, axis=-1, keepdims=True)
    return m * x / r2 ** 1.5


def test_derivs23_are_time_derivatives_of_the_pair_forces():
    # For every pair, the relative orbit x(t) = x0 + v t + A t^2/2 + J t^3/6
    # with A = a_j - a_i and J = adot_j - adot_i has pair force
    # F(t) = m_j x / |x|^3; the start-up terms must be F''(0) and F'''(0).
    # Checked by 4th-order central differences of F, summed over j, with the
    # regular / irregular split by r_nb.
    rng = np.random.default_rng(71)
    n = 7
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3)) * 0.5
    a = rng.normal(size=(n, 3)) * 0.3
    ad = rng.normal(size=(n, 3)) * 0.3
    m = rng.uniform(0.5, 2.0, n)
    rnb2 = np.full(n, 1.0)
    t = np.arange(n, dtype=np.int32)
    out = oracle.derivs23(x, v, a, ad, m, rnb2, t, ref_signs=False)
    h = 2e-3

    def F(i, j, s):
        d0 = x[j] - x[i]
        dv = v[j] - v[i]
        A = a[j] - a[i]
        J = ad[j] - ad[i]
        return _pair_force(m[j], d0 + dv * s + A * s * s / 2 + J * s ** 3 / 6)

    for i in range(n):
        reg2 = np.zeros(3); reg3 = np.zeros(3); irr2 = np.zeros(3); irr3 = np.zeros(3)
        for j in range(n):
            if j == i:
                continue
            f = {k: F(i, j, k * h) for k in (-3, -2, -1, 0, 1, 2, 3)}
            d2 = (-f[2] + 16 * f[1] - 30 * f[0] + 16 * f[-1] - f[-2]) / (12 * h * h)
            d3 = (-f[3] + 8 * f[2] - 13 * f[1] + 13 * f[-1] - 8 * f[-2] + f[-3]) / (8 * h ** 3)
            if np.sum((x[j] - x[i]) ** 2) > rnb2[i]:
                reg2 += d2; reg3 += d3
            else:
                irr2 += d2; irr3 += d3
        scale2 = np.abs(out[i, [0, 1, 2, 6, 7, 8]]).max() + 1
        scale3 = np.abs(out[i, [3, 4, 5, 9, 10, 11]]).max() + 1
        np.testing.assert_allclose(out[i, 0:3], reg2, rtol=0, atol=1e-6 * scale2)
        np.testing.assert_allclose(out[i, 6:9], irr2, rtol=0, atol=1e-6 * scale2)
        np.testing.assert_allclose(out[i, 3:6], reg3, rtol=0, atol=1e-4 * scale3)
        np.testing.assert_allclose(out[i, 9:12], irr3, rtol=0, atol=1e-4 * scale3)
    # both parts occur
    assert np.abs(out[:, 0:6]).max() > 0 and np.abs(out[:, 6:12]).max() > 0


def test_update_radius_end_points():
    fix, nmax = 100, 1000
    r2 = np.full(5, 2.0)
    got = oracle.update_radius([fix, nmax, fix + 1, 0, fix - 1], r2, fix, nmax)
    # nnb = fix: unchanged; nnb = nmax: a = 0.9 -> x0.1; fix+1: a = 0.9/900^2
    # nnb = 0: a = 0.5 (-100)^3 / 900^3 (< 0: grows); fix-1: a = -0.5/900^3
    np.testing.assert_allclose(got, [2.0, 0.2, 2.0 * (1 - 0.9 / 810000.0),
                                     2.0 * (1 + 0.5e6 / 729e6), 2.0 * (1 + 0.5 / 729e6)],
                               rtol=1e-15)


def test_derivs23_reference_signs():
    # The reference's projections x.(a1 - a2) etc. carry the opposite sign to
    # the relative acceleration its leading terms use.  With equal
    # accelerations and jerks everywhere the projections vanish and both
    # variants equal the exact derivatives; with only the accelerations
    # differing, adot2 differs by exactly -3 (B_ref - B) a21 = 6 (x.A/r2) a21
    # per pair (A = a_j - a_i), checked here on a single pair.
    rng = np.random.default_rng(73)
    x = rng.normal(size=(5, 3)); v = rng.normal(size=(5, 3)); m = rng.uniform(0.5, 2, 5)
    a_same = np.tile(rng.normal(size=3), (5, 1)); ad_same = np.tile(rng.normal(size=3), (5, 1))
    t = np.arange(5, dtype=np.int32)
    r = oracle.derivs23(x, v, a_same, ad_same, m, np.zeros(5), t, ref_signs=True)
    e = oracle.derivs23(x, v, a_same, ad_same, m, np.zeros(5), t, ref_signs=False)
    np.testing.assert_allclose(r, e, rtol=1e-13, atol=1e-13)
    xp = np.array([[0.0, 0, 0], [1.0, 2.0, -0.5]]); vp = np.zeros((2, 3))
    ap = np.array([[0.0, 0, 0], [0.3, -0.2, 0.7]]); adp = np.zeros((2, 3)); mp = np.array([1.0, 2.0])
    r = oracle.derivs23(xp, vp, ap, adp, mp, np.zeros(2), [0], ref_signs=True)[0]
    e = oracle.derivs23(xp, vp, ap, adp, mp, np.zeros(2), [0], ref_signs=False)[0]
    d = xp[1]; r2 = d @ d; A = ap[1] - ap[0]
    a21 = mp[1] * d / r2 ** 1.5
    np.testing.assert_allclose(r[0:3] - e[0:3], 6 * (d @ A) / r2 * a21, rtol=1e-12)
