"""ORACLE -- test infrastructure only (see regf_oracle.c for what it restates).

ctypes wrapper over oracle/liboracle.so.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this package, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        from . import build_oracle
        build_oracle.build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.oracle_regf.restype = ctypes.c_int
    lib.oracle_regf.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _ip,
                                _dp, _dp, _ip, _ip, ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.oracle_band_pairs.restype = None
    lib.oracle_band_pairs.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _ip,
                                      ctypes.c_double, _ip, _ip, ctypes.c_int, ctypes.c_int]
    lib.oracle_predict2.restype = None
    lib.oracle_predict2.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_irregular.restype = None
    lib.oracle_irregular.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, _ip, _ip,
                                     ctypes.c_int, _dp, _dp]
    lib.oracle_regular_correct.restype = None
    lib.oracle_regular_correct.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, _dp,
                                           _dp, _dp, _dp, _ip, _ip, _ip, _ip, ctypes.c_int, _dp]
    lib.oracle_derivs23.restype = None
    lib.oracle_derivs23.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _ip,
                                    ctypes.c_int, ctypes.c_int, _dp]
    lib.oracle_update_radius.restype = None
    lib.oracle_update_radius.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp]
    lib.oracle_max_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def regf(x, v, m, rnb2, targets, cap, nthreads=0, scales=False):
    """FP64 regular force + neighbour lists.

    Returns (acc[nt,3], adot[nt,3], nnb[nt], lists[nt,cap]) and, with
    scales=True, also (sum_j |a-term|, sum_j |jdot-term|) per target.
    nnb[k] < 0 flags overflow (-(true count)); the first cap entries are kept.
    """
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(m, dtype=np.float64)
    rnb2 = np.ascontiguousarray(rnb2, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    n = x.shape[0]
    assert v.shape == (n, 3) and m.shape == (n,) and rnb2.shape == (n,)
    if t.size and (t.min() < 0 or t.max() >= n):
        raise IndexError("target index out of range")
    nt = t.size
    acc = np.zeros((nt, 3))
    adot = np.zeros((nt, 3))
    nnb = np.zeros(nt, dtype=np.int32)
    lists = np.full((nt, max(cap, 1)), -1, dtype=np.int32)
    sa = np.zeros(nt) if scales else None
    sj = np.zeros(nt) if scales else None
    lib.oracle_regf(n, _d(x), _d(v), _d(m), _d(rnb2), nt, _i(t), _d(acc), _d(adot),
                    _i(nnb), _i(lists), cap, _d(sa) if scales else None,
                    _d(sj) if scales else None, nthreads)
    if scales:
        return acc, adot, nnb, lists, sa, sj
    return acc, adot, nnb, lists


def band_pairs(x, rnb2, targets, rel=1e-6, bcap=64, nthreads=0):
    """Per target: j != i with |r^2 - r_nb^2| <= rel * r_nb^2 (count, list)."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    rnb2 = np.ascontiguousarray(rnb2, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nt = t.size
    bcnt = np.zeros(nt, dtype=np.int32)
    blist = np.full((nt, bcap), -1, dtype=np.int32)
    lib.oracle_band_pairs(x.shape[0], _d(x), _d(rnb2), nt, _i(t), rel, _i(bcnt),
                          _i(blist), bcap, nthreads)
    return bcnt, blist


def predict2(x, v, a, adot, dt):
    """Second-order predictor of the reference caller (per-particle dt)."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
    adot = np.ascontiguousarray(adot, dtype=np.float64).reshape(-1, 3)
    dt = np.ascontiguousarray(np.broadcast_to(dt, (x.shape[0],)), dtype=np.float64)
    xp = np.empty_like(x)
    vp = np.empty_like(v)
    lib.oracle_predict2(x.shape[0], _d(x), _d(v), _d(a), _d(adot), _d(dt), _d(xp), _d(vp))
    return xp, vp


def irregular(x, v, m, targets, nnb, lists, cap):
    """FP64 irregular force + jerk of each target over its list row (stride cap)."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(m, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nnb = np.ascontiguousarray(nnb, dtype=np.int32)
    lists = np.ascontiguousarray(lists, dtype=np.int32).reshape(-1)
    assert nnb.size == t.size and lists.size >= t.size * cap
    acc = np.empty((t.size, 3))
    adot = np.empty((t.size, 3))
    lib.oracle_irregular(x.shape[0], _d(x), _d(v), _d(m), t.size, _i(t), _i(nnb), _i(lists),
                         cap, _d(acc), _d(adot))
    return acc, adot


def regular_correct(x, v, m, targets, dt, a_old, new_acc, new_adot, old_nnb, old_lists,
                    new_nnb, new_lists, cap):
    """Neighbour-set correction + regular corrector: out [nt, 18] =
    x_new, v_new, a2, a3, a_irr, adot_irr."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(m, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nt = t.size
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (nt,)))
    a_old = np.ascontiguousarray(a_old, dtype=np.float64).reshape(nt, 6)
    new_acc = np.ascontiguousarray(new_acc, dtype=np.float64).reshape(nt, 3)
    new_adot = np.ascontiguousarray(new_adot, dtype=np.float64).reshape(nt, 3)
    old_nnb = np.ascontiguousarray(old_nnb, dtype=np.int32)
    new_nnb = np.ascontiguousarray(new_nnb, dtype=np.int32)
    old_lists = np.ascontiguousarray(old_lists, dtype=np.int32).reshape(-1)
    new_lists = np.ascontiguousarray(new_lists, dtype=np.int32).reshape(-1)
    assert old_nnb.size == nt and new_nnb.size == nt
    assert old_lists.size >= nt * cap and new_lists.size >= nt * cap
    out = np.empty((nt, 18))
    lib.oracle_regular_correct(x.shape[0], _d(x), _d(v), _d(m), nt, _i(t), _d(dt), _d(a_old),
                               _d(new_acc), _d(new_adot), _i(old_nnb), _i(old_lists),
                               _i(new_nnb), _i(new_lists), cap, _d(out))
    return out


def derivs23(x, v, a, adot, m, rnb2, targets, ref_signs=True, nthreads=0):
    """Start-up 2nd/3rd acceleration derivatives: out [nt, 12] =
    a_reg2, a_reg3, a_irr2, a_irr3 (ref_signs: the reference's projection
    signs; False: the exact derivatives)."""
    lib = _load()
    f = lambda z: np.ascontiguousarray(z, dtype=np.float64).reshape(-1, 3)
    x, v, a, adot = f(x), f(v), f(a), f(adot)
    m = np.ascontiguousarray(m, dtype=np.float64)
    rnb2 = np.ascontiguousarray(rnb2, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    out = np.empty((t.size, 12))
    lib.oracle_derivs23(x.shape[0], _d(x), _d(v), _d(a), _d(adot), _d(m), _d(rnb2), t.size,
                        _i(t), 1 if ref_signs else 0, nthreads, _d(out))
    return out


def update_radius(nnb, r2, fix=100, nmax=1000):
    """Neighbour-radius adaptation of r2 (squared radius) from the counts."""
    lib = _load()
    nnb = np.ascontiguousarray(nnb, dtype=np.int32)
    r2 = np.array(r2, dtype=np.float64)
    lib.oracle_update_radius(nnb.size, _i(nnb), fix, nmax, _d(r2))
    return r2


def max_threads():
    return _load().oracle_max_threads()
