"""Seeded initial-condition generators for the regular-force workloads.

All outputs are FP64 in Henon units (G = M = 1, E = -1/4) unless stated:
positions x[n, 3], velocities v[n, 3], masses m[n].

* Plummer sphere: Aarseth, Henon & Wielen (1974) sampling, scale a = 3*pi/16.
* King (1966) model of central depth W0: the isothermal-lowered DF
  f(E) ~ exp(-E/s^2) - 1, Poisson equation integrated in King units, radii by
  inverse CDF of M(r), speeds by rejection from v^2 (exp(W - v^2/2) - 1).
* Kroupa (2001) IMF: dN/dm ~ m^-1.3 on [0.08, 0.5], m^-2.3 on [0.5, 100] Msun,
  normalised to total mass 1.
* Hard binaries: disjoint random pairs turned into circular binaries about the
  primary's position, semi-major axis log-uniform, random orientation.
* Neighbour radius: r_nb^2 per particle = midpoint of the k-th and (k+1)-th
  nearest-neighbour r^2, so about k neighbours and no pair on the boundary.

The reference reads its initial conditions from text files
(src_mpi/ReadWrite.cpp:20-107); these synthetic generators replace that input
for the regular-force workloads.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["plummer", "king", "kroupa_masses", "add_hard_binaries",
           "neighbour_radius2", "log_uniform_block_sizes", "to_com_frame", "energy"]


def _iso_dirs(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def to_com_frame(x, v, m):
    """Shift to the centre-of-mass frame (positions and velocities)."""
    w = m / m.sum()
    return x - (w[:, None] * x).sum(0), v - (w[:, None] * v).sum(0)


def kroupa_masses(n, rng, mmin=0.08, mbreak=0.5, mmax=100.0, a1=1.3, a2=2.3):
    """Kroupa (2001) two-segment IMF, masses normalised to sum to 1."""
    # segment integrals of m^-a dm with continuity at mbreak
    def seg(a, lo, hi):
        return (hi ** (1 - a) - lo ** (1 - a)) / (1 - a)
    c2 = mbreak ** (a2 - a1)          # k2 / k1 so that k1 mb^-a1 == k2 mb^-a2
    w1 = seg(a1, mmin, mbreak)
    w2 = c2 * seg(a2, mbreak, mmax)
    u = rng.uniform(0.0, 1.0, n)
    first = u < w1 / (w1 + w2)
    out = np.empty(n)
    # inverse CDF per segment
    u1 = rng.uniform(0.0, 1.0, n)
    lo, hi = mmin ** (1 - a1), mbreak ** (1 - a1)
    out[first] = (lo + u1[first] * (hi - lo)) ** (1.0 / (1 - a1))
    lo, hi = mbreak ** (1 - a2), mmax ** (1 - a2)
    out[~first] = (lo + u1[~first] * (hi - lo)) ** (1.0 / (1 - a2))
    return out / out.sum()


def plummer(n, seed=1, masses=None, rmax_frac=0.999):
    """Plummer sphere in Henon units. masses: None -> equal masses 1/n."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n) if masses is None else np.asarray(masses, float) / np.sum(masses)
    a = 3.0 * math.pi / 16.0
    xm = rng.uniform(0.0, rmax_frac, n)                # mass fraction inside r
    r = a / np.sqrt(xm ** (-2.0 / 3.0) - 1.0)
    x = r[:, None] * _iso_dirs(rng, n)
    # speed fraction q = v / v_esc with density q^2 (1 - q^2)^3.5 (rejection)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0) * (r * r + a * a) ** -0.25
    v = (q * vesc)[:, None] * _iso_dirs(rng, n)
    x, v = to_com_frame(x, v, m)
    return x, v, m


def _king_profile(W0, npts=20000):
    """Dimensionless King model: returns r, W(r), M(r), potential energy U.

    Units: G = 1, sigma = 1, King radius r0 = 1 so that rho0 = 9/(4 pi).
    """
    from math import erf, sqrt, pi, exp

    def rho_hat(W):
        if W <= 0.0:
            return 0.0
        sw = sqrt(W)
        return exp(W) * erf(sw) - sqrt(4.0 * W / pi) * (1.0 + 2.0 * W / 3.0)

    rho0_hat = rho_hat(W0)
    rho_c = 9.0 / (4.0 * pi)

    def rho(W):
        return rho_c * rho_hat(W) / rho0_hat

    # RK4 on y = (W, dW/dr, M) in r, series start near the centre
    h = 1e-3
    r = h
    W = W0 - 1.5 * r * r          # W ~ W0 - (4 pi rho_c / 6) r^2 with 4 pi rho_c = 9
    dW = -3.0 * r
    M = 4.0 / 3.0 * pi * r ** 3 * rho_c
    rs, Ws, Ms = [0.0, r], [W0, W], [0.0, M]

    def deriv(rr, y):
        w, dw, mm = y
        rr_ = rho(w)
        return (dw, -4.0 * pi * rr_ - 2.0 * dw / rr, 4.0 * pi * rr * rr * rr_)

    y = (W, dW, M)
    while y[0] > 0.0 and r < 1e4:
        # adaptive step: proportional to r, finer near the edge
        h = min(0.02 * max(r, 0.05), 0.5)
        k1 = deriv(r, y)
        k2 = deriv(r + h / 2, tuple(a + h / 2 * b for a, b in zip(y, k1)))
        k3 = deriv(r + h / 2, tuple(a + h / 2 * b for a, b in zip(y, k2)))
        k4 = deriv(r + h, tuple(a + h * b for a, b in zip(y, k3)))
        ynew = tuple(a + h / 6 * (b + 2 * c + 2 * d + e)
                     for a, b, c, d, e in zip(y, k1, k2, k3, k4))
        if ynew[0] <= 0.0:
            # linear interpolation to the tidal radius
            f = y[0] / (y[0] - ynew[0])
            r_t = r + f * h
            M_t = y[2] + f * (ynew[2] - y[2])
            rs.append(r_t); Ws.append(0.0); Ms.append(M_t)
            break
        r += h
        y = ynew
        rs.append(r); Ws.append(y[0]); Ms.append(y[2])
    rs = np.array(rs); Ws = np.array(Ws); Ms = np.array(Ms)
    # potential energy U = -int G M(r) dM / r
    dM = np.diff(Ms)
    rmid = 0.5 * (rs[1:] + rs[:-1])
    Mmid = 0.5 * (Ms[1:] + Ms[:-1])
    ok = rmid > 0
    U = -np.sum(Mmid[ok] * dM[ok] / rmid[ok])
    return rs, Ws, Ms, U


def king(n, W0=6.0, seed=1, masses=None):
    """King (1966) model of depth W0 in Henon units."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n) if masses is None else np.asarray(masses, float) / np.sum(masses)
    rs, Ws, Ms, U = _king_profile(float(W0))
    Mt = Ms[-1]
    # radii by inverse CDF of M(r)
    u = rng.uniform(0.0, 1.0, n) * Mt
    r = np.interp(u, Ms, rs)
    W = np.interp(r, rs, Ws)
    W = np.maximum(W, 1e-12)
    # speed: y = v / sqrt(2W) in [0,1], density y^2 (exp(W (1 - y^2)) - 1)
    grid = np.linspace(0.0, 1.0, 65)
    speed = np.empty(n)
    todo = np.arange(n)
    chunk = 1 << 16
    gmax = np.empty(n)
    for s in range(0, n, chunk):
        w = W[s:s + chunk, None]
        g = grid[None, :] ** 2 * np.expm1(w * (1.0 - grid[None, :] ** 2))
        gmax[s:s + chunk] = g.max(axis=1) * 1.05 + 1e-300
    while todo.size:
        y = rng.uniform(0.0, 1.0, todo.size)
        t = rng.uniform(0.0, 1.0, todo.size) * gmax[todo]
        w = W[todo]
        ok = t < y * y * np.expm1(w * (1.0 - y * y))
        speed[todo[ok]] = y[ok] * np.sqrt(2.0 * w[ok])
        todo = todo[~ok]
    x = r[:, None] * _iso_dirs(rng, n)
    v = speed[:, None] * _iso_dirs(rng, n)
    # model units (G=1, mass Mt) -> Henon units (M = 1, E = -1/4)
    E_model = U / 2.0 / (Mt * Mt)        # after masses -> /Mt, velocities -> /sqrt(Mt)
    lam = E_model / -0.25
    x = x * lam
    v = v / math.sqrt(Mt) / math.sqrt(lam)
    x, v = to_com_frame(x, v, m)
    return x, v, m


def add_hard_binaries(x, v, m, frac=0.05, amin=1e-5, amax=1e-3, seed=1):
    """Turn `frac` of the particles into circular binaries.

    Disjoint random pairs (p, s) are chosen; the secondary s is moved onto a
    circular orbit about the pair's centre of mass, which stays at p's former
    position with p's former velocity.  Returns new (x, v, m) copies and the
    (p, s) index pairs.
    """
    rng = np.random.default_rng(seed + 7919)
    n = len(m)
    npair = int(round(frac * n / 2))
    idx = rng.permutation(n)[:2 * npair].reshape(npair, 2)
    x = x.copy(); v = v.copy()
    p, s = idx[:, 0], idx[:, 1]
    mp, ms = m[p], m[s]
    mt = mp + ms
    a = np.exp(rng.uniform(math.log(amin), math.log(amax), npair))
    e1 = _iso_dirs(rng, npair)
    # a unit vector perpendicular to e1
    t = _iso_dirs(rng, npair)
    e2 = np.cross(e1, t)
    e2 /= np.linalg.norm(e2, axis=1)[:, None]
    vrel = np.sqrt(mt / a)
    xc, vc = x[p].copy(), v[p].copy()
    x[p] = xc - (ms / mt * a)[:, None] * e1
    x[s] = xc + (mp / mt * a)[:, None] * e1
    v[p] = vc - (ms / mt * vrel)[:, None] * e2
    v[s] = vc + (mp / mt * vrel)[:, None] * e2
    return x, v, m.copy(), idx


def neighbour_radius2(x, k, workers=-1):
    """r_nb^2 per particle: midpoint of the k-th and (k+1)-th neighbour r^2."""
    from scipy.spatial import cKDTree
    tree = cKDTree(x)
    kk = min(k + 2, len(x))
    d, _ = tree.query(x, k=kk, workers=workers)
    d2 = d * d
    # column 0 is the particle itself (distance 0)
    lo = d2[:, min(k, kk - 1)]
    hi = d2[:, min(k + 1, kk - 1)]
    return 0.5 * (lo + hi)


def log_uniform_block_sizes(count, lo, hi, seed=1):
    """Active-block sizes drawn log-uniform on [lo, hi] (integers)."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(math.log(lo), math.log(hi + 1), count))
    return np.clip(s.astype(np.int64), lo, hi)


def energy(x, v, m, chunk=2048):
    """Total energy (kinetic + potential), O(n^2) in chunks; for tests."""
    K = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    U = 0.0
    n = len(m)
    for s in range(0, n, chunk):
        d = x[s:s + chunk, None, :] - x[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=2))
        idx = np.arange(s, min(s + chunk, n))
        r[np.arange(len(idx)), idx] = np.inf
        U -= 0.5 * np.sum(m[s:s + chunk, None] * m[None, :] / r)
    return K + U
