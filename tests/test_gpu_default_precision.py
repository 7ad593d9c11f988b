"""The default precision at the accuracy contract: a and jdot within 1e-6
relative of FP64 on EVERY target, on the three single-cluster workloads,
through a context with default settings and through the six legacy entry
points (the path ABYSS calls, src_mpi/cuda/cuda_functions.h:3-10).

  config 1: equal-mass Plummer N = 1024, all targets;
  config 2: Plummer + Kroupa N = 65536, cap 256, blocks of 1 ... 65536
            targets (log-spaced sizes, random targets);
  config 3: King W0 = 6 + 5 % hard binaries N = 262144, all targets against
            the FP64 kernel and a 1024-target sample against the oracle.

Rows whose regular set the FP64 evaluation does not agree with (a pair in
the 1e-6 band of r_nb^2 falls the other way) are compared through their
lists (no hard mismatch), not their sums.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from oracle.compare import compare_lists, force_errors
from abyss_amd import regf, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-6
BLOCKS = (1, 2, 5, 16, 64, 100, 256, 1000, 4096, 16384, 65536)


def say(*a):
    print(*a, flush=True)


def same_set_rows(nnb, lst, rn, rl):
    w = lst.shape[1]
    valid = np.arange(w)[None, :] < np.minimum(np.abs(nnb), w)[:, None]
    return (nnb == rn) & np.all((lst == rl[:, :w]) | ~valid, axis=1)


def assert_contract(label, js, t, res, ref=None, cap=None):
    """res against the oracle (or against `ref` = (acc, adot, nnb, lst) of
    the FP64 kernel): lists exact outside the band, max error <= TOL."""
    acc, adot, nnb, lst = res
    cap = lst.shape[1] if cap is None else cap
    if ref is None:
        ra, rj, rn, rl = oracle.regf(js.x, js.v, js.m, js.r2, t, cap)
    else:
        ra, rj, rn, rl = ref
    lists = compare_lists(nnb, lst, rn, rl, js.x, js.r2, t)
    rows = same_set_rows(nnb, lst, rn, rl)
    err = force_errors(acc[rows], adot[rows], ra[rows], rj[rows])
    say(f"[{label}] nt={len(t)} band_rows={(~rows).sum()} a_rel_max={err['a_rel']['max']:.3e} "
        f"jdot_rel_max={err['jdot_rel']['max']:.3e} lists={lists}")
    assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0
    assert lists["overflow_mismatch"] == 0
    assert (~rows).sum() <= max(16, len(t) // 10000)
    assert err["a_rel"]["max"] <= TOL, label
    assert err["jdot_rel"]["max"] <= TOL, label
    return err


def test_context_default_is_checked():
    with regf.Context(device=0, cap=64) as c:
        js = workloads.plummer_equal(1024)
        c.upload(js.m, js.x, js.v, js.r2)
        c.query(np.arange(js.n, dtype=np.int32))
        assert c.last_recheck() > 0      # the check ran and repaired targets


def test_config1_context_and_legacy():
    js = workloads.plummer_equal(1024)
    t = np.arange(js.n, dtype=np.int32)
    with regf.Context(device=0, cap=256) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        ctx_res = c.query(t)
    assert_contract("config1/context", js, t, ctx_res)
    regf.OpenDevice(0)
    try:
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        leg = regf.CalculateAccelerationOnDevice(t)
    finally:
        regf.CloseDevice()
    assert_contract("config1/legacy", js, t, leg)
    assert np.array_equal(leg[0], ctx_res[0]) and np.array_equal(leg[1], ctx_res[1])


def test_config2_every_block_size_legacy():
    js = workloads.plummer_kroupa(65536)
    rng = np.random.default_rng(202)
    regf.set_num_neighbor_max(js.cap)
    regf.OpenDevice(0)
    try:
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        for size in BLOCKS:
            say(f"case: config2 legacy block {size}")
            t = rng.permutation(js.n)[:size].astype(np.int32)
            res = regf.CalculateAccelerationOnDevice(t)
            assert_contract(f"config2/legacy/{size}", js, t, res)
    finally:
        regf.CloseDevice()
        regf.set_num_neighbor_max(1000)


def test_config2_every_block_size_context():
    js = workloads.plummer_kroupa(65536)
    rng = np.random.default_rng(203)
    with regf.Context(device=0, cap=js.cap) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        for size in BLOCKS:
            say(f"case: config2 context block {size}")
            t = rng.permutation(js.n)[:size].astype(np.int32)
            assert_contract(f"config2/context/{size}", js, t, c.query(t))


def test_config3_every_target_legacy():
    js = workloads.king_binaries(262144)
    t = np.arange(js.n, dtype=np.int32)
    with regf.Context(device=0, cap=js.cap, precision=regf.PREC_FP64) as c:
        c.upload(js.m, js.x, js.v, js.r2)
        f64 = c.query(t)
    regf.set_num_neighbor_max(js.cap)
    regf.OpenDevice(0)
    try:
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        res = regf.CalculateAccelerationOnDevice(t)
    finally:
        regf.CloseDevice()
        regf.set_num_neighbor_max(1000)
    assert_contract("config3/legacy/all-vs-fp64", js, t, res, ref=f64)
    sample = np.random.default_rng(31).permutation(js.n)[:1024].astype(np.int32)
    assert_contract("config3/legacy/sample-vs-oracle", js, sample,
                    tuple(a[sample] for a in res))


def test_legacy_fp32_abi_library():
    """libabyss_regf_f32.so (the float legacy ABI of an unedited src_mpi
    build) gives the FP64-ABI results rounded to float, for inputs that are
    floats already.  Runs in a child process: the two libraries export the
    same legacy symbols."""
    code = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from abyss_amd import regf, workloads
js = workloads.plummer_kroupa(65536)
x32, v32 = js.x.astype(np.float32), js.v.astype(np.float32)
m32, r32 = js.m.astype(np.float32), js.r2.astype(np.float32)
t = np.random.default_rng(5).permutation(js.n)[:3000].astype(np.int32)
n, nt, cap = js.n, len(t), 256
lib = ctypes.CDLL(regf.LIB_PATH.replace("libabyss_regf.so", "libabyss_regf_f32.so"))
lib.abyss_set_legacy_errors(1)
assert lib.abyss_set_num_neighbor_max(cap) == 0
fp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
ci = lambda v: ctypes.byref(ctypes.c_int(v))
lib.OpenDevice(ci(0))
lib.SendToDevice(ci(n), fp(m32), fp(x32), fp(v32), fp(r32), fp(m32))
acc = np.zeros((nt, 3), np.float32); adot = np.zeros((nt, 3), np.float32)
nnb = np.zeros(nt, np.int32); lst = np.zeros((nt, cap), np.int32)
lib.CalculateAccelerationOnDevice(ci(nt), fp(t), fp(acc), fp(adot), fp(nnb), fp(lst))
assert lib.abyss_last_error() == 0, lib.abyss_last_error()
lib.CloseDevice()
# the FP64 ABI on the same (float-valued) inputs
regf.set_num_neighbor_max(cap)
regf.OpenDevice(0)
regf.SendToDevice(m32.astype(np.float64), x32.astype(np.float64), v32.astype(np.float64),
                  r32.astype(np.float64))
a64, j64, n64, l64 = regf.CalculateAccelerationOnDevice(t)
regf.CloseDevice()
assert np.array_equal(nnb, n64)
filled = np.arange(cap)[None, :] < np.minimum(nnb, cap)[:, None]   # populated list entries
assert np.array_equal(lst[filled], l64[filled])
assert np.array_equal(acc, a64.astype(np.float32)) and np.array_equal(adot, j64.astype(np.float32))
print("f32 ABI ok", nt, flush=True)
'''
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-c", code, root], capture_output=True, text=True,
                       timeout=300)
    say(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0 and "f32 ABI ok" in p.stdout


def test_legacy_default_exits_on_overflow():
    """With the default error mode the legacy entry points stop the process
    on neighbour overflow, as the reference does (exit status 1, message on
    stderr); the caller never sees truncated lists."""
    code = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from abyss_amd import regf, workloads
js = workloads.plummer_equal(1024)
lib = ctypes.CDLL(regf.LIB_PATH)        # the C defaults, no Python mirror
assert lib.abyss_set_num_neighbor_max(8) == 0
fp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
ci = lambda v: ctypes.byref(ctypes.c_int(v))
lib.OpenDevice(ci(0))
x, v, m, r2 = (np.ascontiguousarray(a, np.float64) for a in (js.x, js.v, js.m, js.r2))
lib.SendToDevice(ci(js.n), fp(m), fp(x), fp(v), fp(r2), fp(m))
print("uploaded", flush=True)
t = np.arange(16, dtype=np.int32)
acc = np.zeros((16, 3)); adot = np.zeros((16, 3))
nnb = np.zeros(16, np.int32); lst = np.zeros((16, 8), np.int32)
lib.CalculateAccelerationOnDevice(ci(16), fp(t), fp(acc), fp(adot), fp(nnb), fp(lst))
print("returned", flush=True)
'''
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-c", code, root], capture_output=True, text=True,
                       timeout=300)
    say(p.stdout[-2000:], p.stderr[-2000:])
    assert "uploaded" in p.stdout and "returned" not in p.stdout
    assert p.returncode == 1 and "neighbour list overflow" in p.stderr
