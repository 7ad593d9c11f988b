"""Python mirror of the regular-force C ABI (include/abyss_regf.h).

The six legacy functions keep the reference's names and argument meaning
(src_mpi/cuda/cuda_functions.h:3-10): SendToDevice replaces the j-set,
CalculateAccelerationOnDevice returns the regular acc/adot and ascending
neighbour lists of the targets.  Context wraps the extended per-cluster API.

This module only binds libabyss_regf.so; there is no CPU fallback.  If the
library is missing, loading raises immediately.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libabyss_regf.so")

OK = 0
ERR_NOT_OPEN = 1
ERR_NO_DEVICE = 2
ERR_INVALID_ARG = 3
ERR_TARGET_RANGE = 4
ERR_INVALID_INPUT = 5
ERR_NO_JSET = 6
ERR_HIP = 7
ERR_OUT_OF_MEMORY = 8
ERR_NEIGHBOR_OVERFLOW = 9
ERR_ADMISSION = 10
ERR_BUSY = 11
ERR_SCRATCH_LIMIT = 12
# static device scratch of abyss_query_noalloc, and its j-split share
NOALLOC_SCRATCH_BYTES = 128 << 20
NOALLOC_SPLIT_BYTES = 64 << 20

PREC_SPLIT = 0
PREC_FP64 = 1
PREC_MIXED = 2
PREC_REFINED = 3
PREC_CHECKED = 4
# the library's default precision (contexts, ensembles, servers and the
# legacy entry points): every target's a and adot within 1e-6 of FP64
PREC_DEFAULT = PREC_CHECKED
# abyss_set_legacy_errors modes
LEGACY_EXIT = 0
LEGACY_RETURN = 1
# ABYSS_CHECK_ERR_SCALE: the checked precision recomputes a target when
# CHECK_ERR_SCALE * spread > tol * |a| (or |adot|) for either spread
CHECK_ERR_SCALE = 1.0e-6

# Every exported symbol of include/abyss_regf.h (checked by the CPU tests).
EXPORTS = [
    "InitializeDevice", "OpenDevice", "CloseDevice", "ProfileDevice", "SendToDevice",
    "CalculateAccelerationOnDevice",
    "abyss_last_error", "abyss_error_string", "abyss_last_error_detail",
    "abyss_num_neighbor_max", "abyss_set_num_neighbor_max", "abyss_set_precision",
    "abyss_device_count",
    "abyss_ctx_create", "abyss_ctx_destroy", "abyss_ctx_set_precision", "abyss_ctx_set_tuning",
    "abyss_ctx_set_scratch_limit", "abyss_ctx_upload", "abyss_ctx_query",
    "abyss_ctx_set_targets", "abyss_ctx_launch", "abyss_ctx_sync", "abyss_ctx_fetch",
    "abyss_ctx_stream", "abyss_ctx_plan", "abyss_ctx_error_detail",
    "abyss_ens_create", "abyss_ens_destroy", "abyss_ens_upload", "abyss_ens_release",
    "abyss_ens_set_targets", "abyss_ens_launch", "abyss_ens_sync", "abyss_ens_fetch",
    "abyss_ens_stream", "abyss_ens_device_bytes", "abyss_ens_error_detail",
    "abyss_srv_create", "abyss_srv_destroy", "abyss_srv_add_cluster", "abyss_srv_remove_cluster",
    "abyss_srv_submit", "abyss_srv_wait", "abyss_srv_stats", "abyss_srv_reset_stats",
    "abyss_srv_cluster_footprint", "abyss_srv_set_batch_budget", "abyss_ens_blocks",
    "abyss_ens_set_block_range", "abyss_autotune", "abyss_autotune_load", "abyss_autotune_clear",
    "abyss_autotune_jset", "abyss_ctx_set_shape", "abyss_ctx_plan_block", "abyss_ens_plan",
    "abyss_ctx_upload_state", "abyss_ctx_update_state", "abyss_ctx_predict",
    "abyss_ctx_keep_lists", "abyss_ctx_irregular", "abyss_ctx_regular_step",
    "abyss_ctx_derivs23", "abyss_ctx_update_radius", "abyss_srv_set_admission",
    "abyss_srv_queue_stats", "abyss_ctx_set_check_tolerance", "abyss_ctx_last_recheck",
    "abyss_ctx_fetch_spread", "abyss_ens_set_precision", "abyss_ens_set_check_tolerance",
    "abyss_ens_last_recheck", "abyss_srv_set_precision", "abyss_srv_set_check_tolerance",
    "abyss_set_legacy_errors", "abyss_query_noalloc", "abyss_ctx_set_kernel_timing",
    "abyss_ctx_kernel_timing", "abyss_srv_gpu_busy",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


class AbyssError(RuntimeError):
    def __init__(self, code, detail=""):
        self.code = code
        super().__init__(f"{error_string(code)} (code {code}){': ' + detail if detail else ''}")


def load():
    """Load libabyss_regf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with abyss_amd/build.py "
                          "(no CPU fallback exists)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "InitializeDevice": (None, [_ip]),
        "OpenDevice": (None, [_ip]),
        "CloseDevice": (None, []),
        "ProfileDevice": (None, [_ip]),
        "SendToDevice": (None, [_ip, _dp, _dp, _dp, _dp, _dp]),
        "CalculateAccelerationOnDevice": (None, [_ip, _ip, _dp, _dp, _ip, _ip]),
        "abyss_last_error": (ctypes.c_int, []),
        "abyss_error_string": (ctypes.c_char_p, [ctypes.c_int]),
        "abyss_last_error_detail": (ctypes.c_char_p, []),
        "abyss_num_neighbor_max": (ctypes.c_int, []),
        "abyss_set_num_neighbor_max": (ctypes.c_int, [ctypes.c_int]),
        "abyss_set_precision": (ctypes.c_int, [ctypes.c_int]),
        "abyss_set_legacy_errors": (ctypes.c_int, [ctypes.c_int]),
        "abyss_device_count": (ctypes.c_int, []),
        "abyss_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
        "abyss_ctx_destroy": (ctypes.c_int, [_vp]),
        "abyss_ctx_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_ctx_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int]),
        "abyss_ctx_set_scratch_limit": (ctypes.c_int, [_vp, ctypes.c_ulonglong]),
        "abyss_ctx_upload": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "abyss_ctx_query": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _ip, _ip]),
        "abyss_ctx_set_targets": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
        "abyss_ctx_launch": (ctypes.c_int, [_vp]),
        "abyss_ctx_sync": (ctypes.c_int, [_vp]),
        "abyss_ctx_fetch": (ctypes.c_int, [_vp, _dp, _dp, _ip, _ip]),
        "abyss_ctx_stream": (_vp, [_vp]),
        "abyss_ctx_plan": (ctypes.c_int, [_vp, _ip, _ip, _ip, _ip, _ip, _ip]),
        "abyss_ctx_plan_block": (ctypes.c_int, [_vp, _ip, _ip]),
        "abyss_ens_plan": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
        "abyss_ctx_set_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "abyss_autotune_jset": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                               ctypes.c_char_p]),
        "abyss_ctx_error_detail": (ctypes.c_char_p, [_vp]),
        "abyss_ctx_set_kernel_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_srv_gpu_busy": (ctypes.c_int, [_vp, _dp, _dp]),
        "abyss_ctx_kernel_timing": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
        "abyss_ctx_set_check_tolerance": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_query_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                               ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp,
                                               _vp]),
        "abyss_ctx_last_recheck": (ctypes.c_int, [_vp, _ip]),
        "abyss_ctx_fetch_spread": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
        "abyss_ens_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_ens_set_check_tolerance": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_ens_last_recheck": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
        "abyss_srv_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_srv_set_check_tolerance": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_ens_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(_vp)]),
        "abyss_ens_destroy": (ctypes.c_int, [_vp]),
        "abyss_ens_upload": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "abyss_ens_release": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_ens_set_targets": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip]),
        "abyss_ens_launch": (ctypes.c_int, [_vp]),
        "abyss_ens_sync": (ctypes.c_int, [_vp]),
        "abyss_ens_fetch": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _ip, _ip]),
        "abyss_ens_stream": (_vp, [_vp]),
        "abyss_ens_device_bytes": (ctypes.c_ulonglong, [_vp]),
        "abyss_ens_error_detail": (ctypes.c_char_p, [_vp]),
        "abyss_srv_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_ulonglong, ctypes.POINTER(_vp)]),
        "abyss_srv_destroy": (ctypes.c_int, [_vp]),
        "abyss_srv_add_cluster": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]),
        "abyss_srv_remove_cluster": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_srv_submit": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip,
                                            ctypes.POINTER(ctypes.c_longlong)]),
        "abyss_srv_wait": (ctypes.c_int, [_vp, ctypes.c_longlong, _dp, _dp, _ip, _ip,
                                          ctypes.POINTER(ctypes.c_double)]),
        "abyss_srv_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong),
                                           ctypes.POINTER(ctypes.c_longlong),
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]),
        "abyss_srv_reset_stats": (ctypes.c_int, [_vp]),
        "abyss_srv_set_admission": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_srv_queue_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong),
                                                 ctypes.POINTER(ctypes.c_longlong),
                                                 ctypes.POINTER(ctypes.c_double)]),
        "abyss_srv_cluster_footprint": (ctypes.c_ulonglong, [_vp, ctypes.c_int]),
        "abyss_srv_set_batch_budget": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_ens_blocks": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
        "abyss_ens_set_block_range": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int]),
        "abyss_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
        "abyss_autotune_load": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p]),
        "abyss_autotune_clear": (ctypes.c_int, [ctypes.c_int]),
        "abyss_ctx_upload_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                                  _dp]),
        "abyss_ctx_update_state": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                                                  _dp, _dp]),
        "abyss_ctx_predict": (ctypes.c_int, [_vp, ctypes.c_double]),
        "abyss_ctx_keep_lists": (ctypes.c_int, [_vp, ctypes.c_int]),
        "abyss_ctx_irregular": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp]),
        "abyss_ctx_regular_step": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp,
                                                  _ip, _ip, _dp]),
        "abyss_ctx_derivs23": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_int, _dp]),
        "abyss_ctx_update_radius": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    # The C default ends the process on a failing legacy call, as the
    # reference does; this mirror raises AbyssError instead.
    lib.abyss_set_legacy_errors(LEGACY_RETURN)
    _lib = lib
    return lib


def error_string(code):
    return load().abyss_error_string(code).decode()


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


# --------------------------------------------------------------------------
# Legacy API (reference names)
# --------------------------------------------------------------------------
def last_error():
    return load().abyss_last_error()


def last_error_detail():
    return load().abyss_last_error_detail().decode()


def _check_legacy():
    rc = last_error()
    if rc not in (OK, ERR_NEIGHBOR_OVERFLOW):
        raise AbyssError(rc, last_error_detail())
    return rc


def InitializeDevice(irank=0):
    load().InitializeDevice(ctypes.byref(ctypes.c_int(irank)))
    return _check_legacy()


def OpenDevice(irank=0):
    load().OpenDevice(ctypes.byref(ctypes.c_int(irank)))
    return _check_legacy()


def CloseDevice():
    load().CloseDevice()
    return _check_legacy()


def ProfileDevice(irank=0):
    load().ProfileDevice(ctypes.byref(ctypes.c_int(irank)))
    return last_error()


def SendToDevice(m, x, v, r2, mdot=None):
    """Replace the j-set (x, v: [N,3]; m, r2: [N]); mdot is ignored."""
    x = _f64(x, (-1, 3))
    n = x.shape[0]
    v = _f64(v, (n, 3))
    m = _f64(m, (n,))
    r2 = _f64(r2, (n,))
    md = np.zeros(n) if mdot is None else _f64(mdot, (n,))
    load().SendToDevice(ctypes.byref(ctypes.c_int(n)), _d(m), _d(x), _d(v), _d(r2), _d(md))
    return _check_legacy()


def num_neighbor_max():
    return load().abyss_num_neighbor_max()


def set_num_neighbor_max(cap):
    rc = load().abyss_set_num_neighbor_max(int(cap))
    if rc != OK:
        raise AbyssError(rc, "the stride can only change while the library is closed")


def set_precision(mode):
    rc = load().abyss_set_precision(int(mode))
    if rc != OK:
        raise AbyssError(rc)


def set_legacy_errors(mode):
    """LEGACY_EXIT: a failing legacy call ends the process (the C default);
    LEGACY_RETURN: it records the status (this module's default, which
    turns it into AbyssError)."""
    rc = load().abyss_set_legacy_errors(int(mode))
    if rc != OK:
        raise AbyssError(rc)


def CalculateAccelerationOnDevice(targets, out=None):
    """Returns (acc[nt,3], adot[nt,3], NumNeighbor[nt], NeighborList[nt, NumNeighborMax]).
    `out`: optional preallocated (acc, adot, nnb, lst) of at least those
    shapes (C-contiguous), written in place like the reference caller's
    buffers."""
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nt = t.size
    cap = num_neighbor_max()
    if out is not None:
        acc, adot, nnb, lst = out
        assert acc.shape[0] >= nt and lst.shape[1] == cap and lst.flags.c_contiguous
    else:
        acc = np.zeros((max(nt, 1), 3))
        adot = np.zeros((max(nt, 1), 3))
        nnb = np.zeros(max(nt, 1), dtype=np.int32)
        lst = np.full((max(nt, 1), cap), -1, dtype=np.int32)
    load().CalculateAccelerationOnDevice(ctypes.byref(ctypes.c_int(nt)), _i(t) if nt else None,
                                         _d(acc), _d(adot), _i(nnb), _i(lst))
    _check_legacy()
    return acc[:nt], adot[:nt], nnb[:nt], lst[:nt]


def device_count():
    return load().abyss_device_count()


def autotune(device=0, n=65536, cache_path=None):
    """Measure and install the plan table for `device` (see abyss_autotune)."""
    rc = load().abyss_autotune(device, n, cache_path.encode() if cache_path else None)
    if rc != OK:
        raise AbyssError(rc)


def autotune_jset(device, m, x, v, r2, cache_path=None):
    """Measure and install the plan table on the caller's own j-set (see
    abyss_autotune_jset): m, r2 as [n], x, v as [n][3]."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    n = len(m)
    if x.shape[0] != n or v.shape[0] != n or len(r2) != n:
        raise ValueError("m, x, v and r2 must describe the same particles")
    rc = load().abyss_autotune_jset(device, n, _d(m), _d(x), _d(v), _d(r2),
                                    cache_path.encode() if cache_path else None)
    if rc != OK:
        raise AbyssError(rc)


def autotune_load(device, cache_path):
    rc = load().abyss_autotune_load(device, cache_path.encode())
    if rc != OK:
        raise AbyssError(rc)


def autotune_clear(device=0):
    rc = load().abyss_autotune_clear(device)
    if rc != OK:
        raise AbyssError(rc)


# --------------------------------------------------------------------------
# Context API
# --------------------------------------------------------------------------
class Context:
    """One device-resident j-set (a cluster) with its own stream and buffers."""

    def __init__(self, device=0, cap=256, precision=None):
        """precision: an ABYSS_PREC_* mode, or None for the library default
        (PREC_DEFAULT)."""
        self._lib = load()
        h = _vp()
        rc = self._lib.abyss_ctx_create(device, cap, ctypes.byref(h))
        if rc != OK:
            raise AbyssError(rc)
        self.h = h
        self.cap = cap
        self.n = 0
        self.nt = 0
        if precision is not None:
            self.set_precision(precision)

    def _chk(self, rc, allow_overflow=False):
        if rc == OK or (allow_overflow and rc == ERR_NEIGHBOR_OVERFLOW):
            return rc
        raise AbyssError(rc, self._lib.abyss_ctx_error_detail(self.h).decode())

    def close(self):
        if self.h:
            self._lib.abyss_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, mode):
        self._chk(self._lib.abyss_ctx_set_precision(self.h, int(mode)))

    def set_check_tolerance(self, tol):
        """Relative jdot tolerance of the checked precision (default 1e-6)."""
        self._chk(self._lib.abyss_ctx_set_check_tolerance(self.h, float(tol)))

    def last_recheck(self):
        """Targets the last checked-precision launch recomputed."""
        n = ctypes.c_int(0)
        self._chk(self._lib.abyss_ctx_last_recheck(self.h, ctypes.byref(n)))
        return n.value

    def spread(self):
        """Per target of the last checked-precision launch: the root sums of
        squares of its FP32 run partials of a and of jdot (float32 [nt, 2])."""
        out = np.empty((max(self.nt, 1), 2), dtype=np.float32)
        self._chk(self._lib.abyss_ctx_fetch_spread(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out[:self.nt]

    def set_tuning(self, i_per_lane=0, j_splits=0, j_unroll=0, lanes_per_target=0):
        self._chk(self._lib.abyss_ctx_set_tuning(self.h, int(i_per_lane), int(j_splits),
                                                 int(j_unroll), int(lanes_per_target)))

    def set_shape(self, tile=0, waves_per_block=0):
        """Force the j-tile (64/128/256) and wave64s per workgroup (1/2/4) of
        the lanes-per-target kernels; 0 = automatic."""
        self._chk(self._lib.abyss_ctx_set_shape(self.h, int(tile), int(waves_per_block)))

    def set_scratch_limit(self, nbytes):
        self._chk(self._lib.abyss_ctx_set_scratch_limit(self.h, int(nbytes)))

    def kernel_timing(self, enable):
        """Record HIP events around the force kernel of every launch (and
        reset the totals)."""
        self._chk(self._lib.abyss_ctx_set_kernel_timing(self.h, int(bool(enable))))

    def read_kernel_timing(self):
        """{"force_ms": summed force-kernel time, "launches": count} since
        kernel_timing was last switched."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self._lib.abyss_ctx_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return {"force_ms": ms.value, "launches": n.value}

    def upload(self, m, x, v, r2):
        x = _f64(x, (-1, 3))
        n = x.shape[0]
        self._chk(self._lib.abyss_ctx_upload(self.h, n, _d(_f64(m, (n,))), _d(x),
                                             _d(_f64(v, (n, 3))), _d(_f64(r2, (n,)))))
        self.n = n

    def set_targets(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        self._chk(self._lib.abyss_ctx_set_targets(self.h, t.size, _i(t) if t.size else None))
        self.nt = t.size

    def upload_state(self, m, x, v, a, adot, t, r2):
        """Device-resident state (positions at each particle's own time t)."""
        x = _f64(x, (-1, 3))
        n = x.shape[0]
        self._chk(self._lib.abyss_ctx_upload_state(
            self.h, n, _d(_f64(m, (n,))), _d(x), _d(_f64(v, (n, 3))), _d(_f64(a, (n, 3))),
            _d(_f64(adot, (n, 3))), _d(_f64(t, (n,))), _d(_f64(r2, (n,)))))
        self.n = n

    def update_state(self, idx, m, x, v, a, adot, t, r2):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        k = i.size
        self._chk(self._lib.abyss_ctx_update_state(
            self.h, k, _i(i), _d(_f64(m, (k,))), _d(_f64(x, (k, 3))), _d(_f64(v, (k, 3))),
            _d(_f64(a, (k, 3))), _d(_f64(adot, (k, 3))), _d(_f64(t, (k,))),
            _d(_f64(r2, (k,)))))

    def predict(self, t_new):
        self._chk(self._lib.abyss_ctx_predict(self.h, float(t_new)))

    def keep_lists(self, enable=True):
        """Keep every launch's neighbour lists resident, by particle index."""
        self._chk(self._lib.abyss_ctx_keep_lists(self.h, 1 if enable else 0))

    def irregular(self, targets):
        """Irregular force and jerk over the resident lists: (acc, adot) [nt,3]."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        acc = np.empty((t.size, 3))
        adot = np.empty((t.size, 3))
        self._chk(self._lib.abyss_ctx_irregular(self.h, t.size, _i(t) if t.size else None,
                                                _d(acc), _d(adot)))
        return acc, adot

    def derivs23(self, targets, reference_signs=True):
        """Start-up a2/a3 split: [nt, 12] = a_reg2, a_reg3, a_irr2, a_irr3."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        out = np.empty((max(t.size, 1), 12))
        self._chk(self._lib.abyss_ctx_derivs23(self.h, t.size, _i(t) if t.size else None,
                                               1 if reference_signs else 0, _d(out)))
        return out[:t.size]

    def update_radius(self, fix=100, nmax=1000):
        """Radius rule for the last launch's targets; returns their new r2."""
        out = np.empty(max(self.nt, 1))
        self._chk(self._lib.abyss_ctx_update_radius(self.h, fix, nmax, _d(out)))
        return out[:self.nt]

    def regular_step(self, targets, dt, a_old):
        """Regular step with neighbour-set correction and regular corrector:
        (acc, adot, nnb, lists[nt, cap], out[nt, 18])."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        nt = t.size
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (nt,)))
        a_old = _f64(a_old, (nt, 6))
        acc = np.zeros((max(nt, 1), 3))
        adot = np.zeros((max(nt, 1), 3))
        nnb = np.zeros(max(nt, 1), dtype=np.int32)
        lst = np.full((max(nt, 1), self.cap), -1, dtype=np.int32)
        out = np.zeros((max(nt, 1), 18))
        self.last_status = self._chk(
            self._lib.abyss_ctx_regular_step(self.h, nt, _i(t) if nt else None, _d(dt),
                                             _d(a_old), _d(acc), _d(adot), _i(nnb), _i(lst),
                                             _d(out)), allow_overflow=True)
        self.nt = nt
        return acc[:nt], adot[:nt], nnb[:nt], lst[:nt], out[:nt]

    def launch(self):
        self._chk(self._lib.abyss_ctx_launch(self.h))

    def sync(self):
        self._chk(self._lib.abyss_ctx_sync(self.h))

    def fetch(self):
        nt = self.nt
        acc = np.zeros((nt, 3))
        adot = np.zeros((nt, 3))
        nnb = np.zeros(nt, dtype=np.int32)
        lst = np.full((nt, self.cap), -1, dtype=np.int32)
        self.last_status = self._chk(
            self._lib.abyss_ctx_fetch(self.h, _d(acc), _d(adot), _i(nnb), _i(lst)),
            allow_overflow=True)
        return acc, adot, nnb, lst

    def query(self, targets, out=None):
        """Synchronous query: (acc, adot, nnb, lists[nt, cap]).  `out` may hold
        preallocated (acc[nt,3], adot[nt,3], nnb[nt], lists[nt,cap]) arrays
        (C-contiguous float64 / int32), which are filled and returned."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        nt = t.size
        if out is not None:
            acc, adot, nnb, lst = out
            if (acc.shape != (nt, 3) or adot.shape != (nt, 3) or nnb.shape != (nt,)
                    or lst.shape != (nt, self.cap) or acc.dtype != np.float64
                    or adot.dtype != np.float64 or nnb.dtype != np.int32
                    or lst.dtype != np.int32 or not all(a.flags.c_contiguous for a in out)):
                raise ValueError("query: out arrays must be C-contiguous acc/adot float64 "
                                 "[nt,3], nnb int32 [nt], lists int32 [nt,cap]")
        else:
            acc = np.zeros((max(nt, 1), 3))
            adot = np.zeros((max(nt, 1), 3))
            nnb = np.zeros(max(nt, 1), dtype=np.int32)
            lst = np.full((max(nt, 1), self.cap), -1, dtype=np.int32)
        self.last_status = self._chk(
            self._lib.abyss_ctx_query(self.h, nt, _i(t) if nt else None, _d(acc), _d(adot),
                                      _i(nnb), _i(lst)), allow_overflow=True)
        self.nt = nt
        if out is not None:
            return acc, adot, nnb, lst
        return acc[:nt], adot[:nt], nnb[:nt], lst[:nt]

    def plan(self):
        a, b, c, d, e, f = (ctypes.c_int() for _ in range(6))
        self._chk(self._lib.abyss_ctx_plan(self.h, ctypes.byref(a), ctypes.byref(b),
                                           ctypes.byref(c), ctypes.byref(d), ctypes.byref(e),
                                           ctypes.byref(f)))
        w, t = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.abyss_ctx_plan_block(self.h, ctypes.byref(w), ctypes.byref(t)))
        return {"i_per_lane": a.value, "j_splits": b.value, "grid_x": c.value, "tile": d.value,
                "j_unroll": e.value, "lanes_per_target": f.value, "waves_per_block": w.value,
                "targets_per_block": t.value}

    @property
    def stream(self):
        return self._lib.abyss_ctx_stream(self.h)


# --------------------------------------------------------------------------
# Batched ensemble and served ensemble
# --------------------------------------------------------------------------
class Ensemble:
    """Up to max_clusters device-resident clusters evaluated by one batched
    launch (include/abyss_regf.h, Part 3)."""

    def __init__(self, device=0, cap=256, max_clusters=256):
        self._lib = load()
        h = _vp()
        rc = self._lib.abyss_ens_create(device, cap, max_clusters, ctypes.byref(h))
        if rc != OK:
            raise AbyssError(rc)
        self.h = h
        self.cap = cap
        self.max_clusters = max_clusters
        self.nt = {}

    def _chk(self, rc, allow_overflow=True):
        if rc == OK or (allow_overflow and rc == ERR_NEIGHBOR_OVERFLOW):
            return rc
        raise AbyssError(rc, self._lib.abyss_ens_error_detail(self.h).decode())

    def close(self):
        if self.h:
            self._lib.abyss_ens_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, mode):
        """PREC_CHECKED (default) or PREC_SPLIT for every cluster."""
        self._chk(self._lib.abyss_ens_set_precision(self.h, int(mode)))

    def set_check_tolerance(self, tol):
        self._chk(self._lib.abyss_ens_set_check_tolerance(self.h, float(tol)))

    def last_recheck(self, cluster):
        """Targets of `cluster` the last checked launch recomputed."""
        n = ctypes.c_int(0)
        self._chk(self._lib.abyss_ens_last_recheck(self.h, int(cluster), ctypes.byref(n)))
        return n.value

    def upload(self, cluster, m, x, v, r2):
        x = _f64(x, (-1, 3))
        n = x.shape[0]
        self._chk(self._lib.abyss_ens_upload(self.h, cluster, n, _d(_f64(m, (n,))), _d(x),
                                             _d(_f64(v, (n, 3))), _d(_f64(r2, (n,)))))

    def release(self, cluster):
        self._chk(self._lib.abyss_ens_release(self.h, cluster))
        self.nt.pop(cluster, None)

    def set_targets(self, cluster, targets):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        self._chk(self._lib.abyss_ens_set_targets(self.h, cluster, t.size,
                                                  _i(t) if t.size else None))
        if t.size:
            self.nt[cluster] = t.size
        else:
            self.nt.pop(cluster, None)

    def blocks(self, cluster):
        nb, tpb = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.abyss_ens_blocks(self.h, cluster, ctypes.byref(nb), ctypes.byref(tpb)))
        return nb.value, tpb.value

    def plan(self, cluster):
        """(lanes per target, j per LDS tile) of a cluster's work items."""
        jl, tile = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.abyss_ens_plan(self.h, cluster, ctypes.byref(jl), ctypes.byref(tile)))
        return jl.value, tile.value

    def set_block_range(self, cluster, first, count):
        self._chk(self._lib.abyss_ens_set_block_range(self.h, cluster, first, count))

    def launch(self):
        self._chk(self._lib.abyss_ens_launch(self.h))

    def sync(self):
        self._chk(self._lib.abyss_ens_sync(self.h))

    def fetch(self, cluster, fill=True, out=None):
        """Results of `cluster`; list entries past each count are -1 with
        fill=True and left uninitialised (faster) with fill=False; `out`:
        preallocated (acc, adot, nnb, lst) written in place."""
        nt = self.nt[cluster]
        if out is not None:
            acc, adot, nnb, lst = out
            assert acc.shape[0] >= nt and lst.shape[1] == self.cap and lst.flags.c_contiguous
        else:
            acc = np.empty((nt, 3))
            adot = np.empty((nt, 3))
            nnb = np.empty(nt, dtype=np.int32)
            lst = (np.full((nt, self.cap), -1, dtype=np.int32) if fill
                   else np.empty((nt, self.cap), dtype=np.int32))
        self._chk(self._lib.abyss_ens_fetch(self.h, cluster, _d(acc), _d(adot), _i(nnb),
                                            _i(lst)))
        return acc, adot, nnb, lst

    @property
    def stream(self):
        return self._lib.abyss_ens_stream(self.h)

    def device_bytes(self):
        return self._lib.abyss_ens_device_bytes(self.h)


class Server:
    """Served ensemble: resident clusters, requests from any thread,
    continuous batching by a native worker thread, admission by HBM budget."""

    def __init__(self, device=0, cap=256, max_clusters=64, hbm_budget=64 << 30):
        self._lib = load()
        h = _vp()
        rc = self._lib.abyss_srv_create(device, cap, max_clusters, int(hbm_budget),
                                        ctypes.byref(h))
        if rc != OK:
            raise AbyssError(rc)
        self.h = h
        self.cap = cap

    def set_precision(self, mode):
        """PREC_CHECKED (default) or PREC_SPLIT, from the next batch on."""
        rc = self._lib.abyss_srv_set_precision(self.h, int(mode))
        if rc != OK:
            raise AbyssError(rc)

    def set_check_tolerance(self, tol):
        rc = self._lib.abyss_srv_set_check_tolerance(self.h, float(tol))
        if rc != OK:
            raise AbyssError(rc)

    def close(self):
        if self.h:
            self._lib.abyss_srv_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, n):
        return self._lib.abyss_srv_cluster_footprint(self.h, n)

    def set_batch_budget(self, pairs):
        rc = self._lib.abyss_srv_set_batch_budget(self.h, float(pairs))
        if rc != OK:
            raise AbyssError(rc)

    def add_cluster(self, m, x, v, r2):
        x = _f64(x, (-1, 3))
        n = x.shape[0]
        cid = ctypes.c_int(-1)
        rc = self._lib.abyss_srv_add_cluster(self.h, n, _d(_f64(m, (n,))), _d(x),
                                             _d(_f64(v, (n, 3))), _d(_f64(r2, (n,))),
                                             ctypes.byref(cid))
        if rc != OK:
            raise AbyssError(rc)
        return cid.value

    def remove_cluster(self, cluster):
        rc = self._lib.abyss_srv_remove_cluster(self.h, cluster)
        if rc != OK:
            raise AbyssError(rc)

    def set_admission(self, max_outstanding_pairs):
        """Limit on submitted, not yet completed work in pair-interactions
        (0: unlimited); requests past it are turned away by submit."""
        rc = self._lib.abyss_srv_set_admission(self.h, float(max_outstanding_pairs))
        if rc != OK:
            raise AbyssError(rc)

    def submit(self, cluster, targets):
        """Queues one regular step; returns (ticket, nt), or (None, nt) when
        admission turns the request away."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        tk = ctypes.c_longlong()
        rc = self._lib.abyss_srv_submit(self.h, cluster, t.size, _i(t), ctypes.byref(tk))
        if rc == ERR_BUSY:
            return None, t.size
        if rc != OK:
            raise AbyssError(rc)
        return tk.value, t.size

    def wait(self, ticket, nt, copy=True):
        """Blocks until the request is done.  Returns (acc, adot, nnb, lists,
        latency_ms); with copy=False only the latency (results discarded)."""
        lat = ctypes.c_double()
        if not copy:
            rc = self._lib.abyss_srv_wait(self.h, ticket, None, None, None, None,
                                          ctypes.byref(lat))
            if rc not in (OK, ERR_NEIGHBOR_OVERFLOW):
                raise AbyssError(rc)
            return lat.value
        acc = np.zeros((nt, 3))
        adot = np.zeros((nt, 3))
        nnb = np.zeros(nt, dtype=np.int32)
        lst = np.full((nt, self.cap), -1, dtype=np.int32)
        rc = self._lib.abyss_srv_wait(self.h, ticket, _d(acc), _d(adot), _i(nnb), _i(lst),
                                      ctypes.byref(lat))
        if rc not in (OK, ERR_NEIGHBOR_OVERFLOW):
            raise AbyssError(rc)
        return acc, adot, nnb, lst, lat.value

    def stats(self):
        c, b = ctypes.c_longlong(), ctypes.c_longlong()
        i, p50, p99 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._lib.abyss_srv_stats(self.h, ctypes.byref(c), ctypes.byref(b), ctypes.byref(i),
                                  ctypes.byref(p50), ctypes.byref(p99))
        rej, mp, peak = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        self._lib.abyss_srv_queue_stats(self.h, ctypes.byref(rej), ctypes.byref(mp),
                                        ctypes.byref(peak))
        busy, wall = ctypes.c_double(), ctypes.c_double()
        self._lib.abyss_srv_gpu_busy(self.h, ctypes.byref(busy), ctypes.byref(wall))
        return {"completed": c.value, "batches": b.value, "interactions": i.value,
                "p50_ms": p50.value, "p99_ms": p99.value, "rejected": rej.value,
                "max_pending": mp.value, "peak_outstanding_pairs": peak.value,
                "gpu_busy_ms": busy.value, "wall_ms": wall.value,
                "gpu_busy_frac": busy.value / wall.value if wall.value > 0 else 0.0}

    def reset_stats(self):
        self._lib.abyss_srv_reset_stats(self.h)
