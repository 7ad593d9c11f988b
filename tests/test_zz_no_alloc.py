"""The no-allocation query (abyss_query_noalloc) through the C ABI, against
the context query it mirrors (abyss_ctx_query on a context given the same
precision and split-scratch limit): counts and lists equal, forces equal bit
for bit under split precision and to FP64 rounding under checked precision
(its recomputation is planned differently), on two calls in a row, the
second smaller.  Device buffers come from the HIP runtime through ctypes."""
import ctypes

import numpy as np
import pytest

from abyss_amd import regf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of libamdhip64 (test-side buffers)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def empty(self, nbytes):
        return self.put(np.zeros(nbytes, np.uint8))

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p,
                                  ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def device_jset(hip, js):
    soa = np.concatenate([js.x[:, 0], js.x[:, 1], js.x[:, 2], js.v[:, 0], js.v[:, 1],
                          js.v[:, 2], js.m, js.r2]).astype(np.float64)
    return hip.put(soa)


def noalloc_query(hip, dj, n, targets, cap, prec):
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nt = t.size
    dt = hip.put(t)
    da, dd = hip.empty(nt * 24), hip.empty(nt * 24)
    dn, dl = hip.empty(nt * 4), hip.empty(nt * cap * 4)
    say(f"noalloc: n={n} nt={nt} cap={cap} prec={prec}")
    rc = regf.load().abyss_query_noalloc(0, n, dj, nt, dt, cap, prec, da, dd, dn, dl, None)
    assert rc == regf.OK, regf.error_string(rc)
    hip.sync()
    return (hip.get(da, (nt, 3), np.float64), hip.get(dd, (nt, 3), np.float64),
            hip.get(dn, nt, np.int32), hip.get(dl, (nt, cap), np.int32))


def context_query(js, targets, cap, prec):
    with regf.Context(device=0, cap=cap) as ctx:
        ctx.set_scratch_limit(regf.NOALLOC_SPLIT_BYTES)
        ctx.set_precision(prec)
        ctx.upload(js.m, js.x, js.v, js.r2)
        return ctx.query(np.ascontiguousarray(targets, dtype=np.int32))


def assert_same(got, ref, cap, prec, rows=None):
    a, d, n, l = got
    ra, rd, rn, rl = ref
    sel = slice(None) if rows is None else rows
    a, d, n, l = a[sel], d[sel], n[sel], l[sel]
    ra, rd, rn, rl = ra[sel], rd[sel], rn[sel], rl[sel]
    assert np.array_equal(n, rn)
    filled = np.arange(cap)[None, :] < np.minimum(np.abs(n), cap)[:, None]
    assert np.array_equal(l[filled], rl[filled])
    if prec == regf.PREC_SPLIT:
        assert np.array_equal(a, ra) and np.array_equal(d, rd)
    else:
        ea = np.linalg.norm(a - ra, axis=1) / np.linalg.norm(ra, axis=1)
        ej = np.linalg.norm(d - rd, axis=1) / np.linalg.norm(rd, axis=1)
        same = np.all(a == ra, axis=1) & np.all(d == rd, axis=1)
        say(f"  max rel diff a {ea.max():.2e} jdot {ej.max():.2e}; bitwise rows "
            f"{same.sum()}/{len(same)}")
        assert ea.max() <= 1e-12 and ej.max() <= 1e-12


@pytest.mark.parametrize("prec", [regf.PREC_SPLIT, regf.PREC_CHECKED])
def test_noalloc_matches_context_two_calls(prec):
    """Plummer + Kroupa N = 65536, cap 256: 4096 targets, then 16."""
    js = workloads.plummer_kroupa(65536)
    rng = np.random.default_rng(11)
    hip = Hip()
    try:
        dj = device_jset(hip, js)
        for nt in (4096, 16):
            t = rng.permutation(js.n)[:nt]
            got = noalloc_query(hip, dj, js.n, t, js.cap, prec)
            ref = context_query(js, t, js.cap, prec)
            assert_same(got, ref, js.cap, prec)
    finally:
        hip.free()


def test_noalloc_full_n_king_checked():
    """King W0 = 6 + binaries N = 262144, every target (one split, checked),
    then 1000 targets."""
    js = workloads.king_binaries(262144)
    rng = np.random.default_rng(12)
    hip = Hip()
    try:
        dj = device_jset(hip, js)
        for t in (np.arange(js.n), rng.permutation(js.n)[:1000]):
            got = noalloc_query(hip, dj, js.n, t, js.cap, regf.PREC_CHECKED)
            ref = context_query(js, t, js.cap, regf.PREC_CHECKED)
            assert_same(got, ref, js.cap, regf.PREC_CHECKED)
    finally:
        hip.free()


def test_noalloc_bad_targets_and_limit():
    """Targets outside [0, n) give NaN forces and INT_MIN counts, the other
    rows equal the context's on the same list with those entries set to 0;
    sizes beyond the static scratch are refused before anything runs."""
    js = workloads.plummer_equal(1024)
    t = np.arange(0, 1024, 8).astype(np.int32)
    bad = np.array([3, 40, 77])
    t[bad] = [-1, 1024, 5000]
    hip = Hip()
    try:
        dj = device_jset(hip, js)
        got = noalloc_query(hip, dj, js.n, t, js.cap, regf.PREC_SPLIT)
        assert np.all(np.isnan(got[0][bad])) and np.all(np.isnan(got[1][bad]))
        assert np.all(got[2][bad] == np.iinfo(np.int32).min)
        t0 = t.copy()
        t0[bad] = 0
        ref = context_query(js, t0, js.cap, regf.PREC_SPLIT)
        good = np.setdiff1d(np.arange(len(t)), bad)
        assert_same(got, ref, js.cap, regf.PREC_SPLIT, rows=good)
        # 40M particles need 160 MB of FP32 masses alone: refused up front
        lib = regf.load()
        dummy = hip.empty(64)
        rc = lib.abyss_query_noalloc(0, 40_000_000, dummy, 4, dummy, 16, regf.PREC_SPLIT,
                                     dummy, dummy, dummy, dummy, None)
        assert rc == regf.ERR_SCRATCH_LIMIT
        rc = lib.abyss_query_noalloc(0, 1024, dj, 4, dummy, 16, regf.PREC_FP64,
                                     dummy, dummy, dummy, dummy, None)
        assert rc == regf.ERR_INVALID_ARG
    finally:
        hip.free()
