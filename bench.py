#!/usr/bin/env python3
"""Benchmark: regular-force pair-interactions/s on one MI355X.

Workload (default): King W0=6 cluster, N=262144, Kroupa IMF, 5% of the
particles in hard binaries, r_nb for ~100 neighbours, list cap 256; one
step = one full-N regular-force + neighbour-list evaluation (m = N targets)
through the library's device-resident launch path (inputs already in HBM),
in the library's default precision (checked: every target's a and jdot
within 1e-6 of FP64; --precision split / fp64 for the others).

Prints ONE JSON line.  Progress goes to stderr.
  value     = interactions / s = steps * m * (N - 1) / (HIP-event time of the
              whole step: force kernel, merge, flagging and recomputation)
  pct_fp32_peak = the step's rate at 60 flop per interaction (the reference's
              convention) against the 157.3 TFLOP/s FP32 vector peak
  roofline  = the force kernel alone: the same flop over its own HIP-event
              time (events around it on the context stream, every timed
              launch); traffic = its HBM bytes per launch from two rocprofv3
              counter passes in this run (FETCH_SIZE, WRITE_SIZE, each scaled
              by its reading for a copy of known size)
  cpu_baseline = the oracle's FP64 C loop (OpenMP over targets) on a bounded
              sample of the same workload on this host's cores; its results on
              256 of those targets also check the GPU results
  other_precisions = the same step in the other of split / checked precision
  accuracy  = per-particle errors over every target against the FP64 kernel
              (split and checked precision), plus the oracle sample
  active_blocks = Plummer + Kroupa N = 65536: blocks of 1, 64, 4096, 65536
              targets (launch, synchronous query, legacy upload + call)
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
# the precision modes a run can name (the library's ABYSS_PREC_* codes)
PREC_NAMES = ("checked", "split", "fp64")


def prec_code(name):
    from abyss_amd import regf
    return {"checked": regf.PREC_CHECKED, "split": regf.PREC_SPLIT,
            "fp64": regf.PREC_FP64}[name]
FLOP_PER_INTERACTION = 60
HBM_PEAK_BPS = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an explicit stream (the library's context stream)."""

    def __init__(self):
        lib = None
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if lib is None:
            raise RuntimeError("libamdhip64.so not found")
        self.lib = lib
        lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                            ctypes.c_void_p]
        lib.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.lib.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return float(ms.value)


def copy_calibration(nbytes, reps=4):
    """Device-to-device copies of a known size (a blit kernel): the reference
    point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    lib = HipEvents().lib
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.hipMalloc(ctypes.byref(a), nbytes) == 0
    assert lib.hipMalloc(ctypes.byref(b), nbytes) == 0
    assert lib.hipMemset(a, 1, nbytes) == 0
    for _ in range(reps):
        assert lib.hipMemcpy(b, a, nbytes, 3) == 0      # hipMemcpyDeviceToDevice
    assert lib.hipDeviceSynchronize() == 0
    lib.hipFree(a)
    lib.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def _counter_rows(outdir):
    """(kernel name, counter value) per dispatch from rocprofv3's CSV files
    under outdir, in dispatch order."""
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"),
                              recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r.get("Dispatch_Id") or 0), r.get("Kernel_Name", ""),
                             float(r.get("Counter_Value") or 0.0)))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def live_traffic(precision, kernel="regf_kernel_jl", timeout=300):
    """HBM bytes of one launch of the force kernel, measured in this run: two
    rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch)
    over `bench.py --traffic-probe`, which copies 256 MiB device to device
    (the reference that scales each counter to bytes) and then runs two
    full-N steps (the second is read).  Returns (info, None) or (None, why)."""
    import shutil
    import signal
    import subprocess
    import tempfile
    if "rocprofiler" in os.environ.get("LD_PRELOAD", ""):
        return None, "bench.py itself runs under a profiler"
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    calib = 256 << 20
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"abyss_pmc_{counter.lower()}_")
        cmd = [exe, "--pmc", counter, "-d", d, "-o", "run", "-f", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe",
               "--precision", precision]
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                             start_new_session=True, cwd=ROOT)
        try:
            _, err = p.communicate(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return None, f"rocprofv3 --pmc {counter} timed out"
        rows = _counter_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        copy = [v for k, v in rows if "copyBuffer" in k]
        force = [v for k, v in rows if kernel in k]
        if p.returncode != 0 and not (copy and force):
            return None, f"rocprofv3 --pmc {counter} failed ({p.returncode}): " + \
                err.decode(errors="replace")[-300:]
        if not copy or not force:
            return None, f"rocprofv3 --pmc {counter}: no copy / force-kernel rows"
        factor = (calib / 1024.0) / (sum(copy) / len(copy))
        out[counter] = {"kib": force[-1], "copy_kib": sum(copy) / len(copy), "factor": factor,
                        "bytes": force[-1] * 1024.0 * factor}
    total = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return {"bytes": total, "fetch_bytes": out["FETCH_SIZE"]["bytes"],
            "write_bytes": out["WRITE_SIZE"]["bytes"],
            "fetch_factor": out["FETCH_SIZE"]["factor"],
            "write_factor": out["WRITE_SIZE"]["factor"],
            "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the "
                      "second full-N launch, each scaled by the same counter's reading for a "
                      "256 MiB device-to-device copy"}, None


def traffic_probe(args):
    """Counter-run body of live_traffic: a 256 MiB copy, then two full-N
    steps of the King N=262144 workload in the given precision."""
    import numpy as np
    from abyss_amd import regf, workloads
    copy_calibration(256 << 20, reps=2)
    js = workloads.king_binaries(args.n)
    ctx = regf.Context(device=0, cap=js.cap, precision=prec_code(args.precision))
    ctx.upload(js.m, js.x, js.v, js.r2)
    ctx.set_targets(np.arange(js.n, dtype=np.int32))
    for _ in range(2):
        ctx.launch()
        ctx.sync()
    ctx.close()


def cpu_baseline(js, target_seconds=10.0, gpu_results=None, nsample=256):
    """Time the oracle's FP64 loop (all host cores) on a bounded target sample.
    With gpu_results ({name: (acc, adot, nnb, lists)} over all targets), the
    oracle's results on the first `nsample` targets of that sample also check
    each GPU result set (returned as the second value)."""
    import numpy as np
    import oracle
    threads = oracle.max_threads()
    rng = np.random.default_rng(7)
    order = rng.permutation(js.n).astype(np.int32)
    # calibrate twice (a tiny probe underestimates the rate: thread start-up),
    # then size the timed sample to ~target_seconds
    size = max(threads * 4, 64)
    rate = 0.0
    for probe_seconds in (None, 1.0):
        if probe_seconds is not None:
            size = int(min(js.n, max(size, rate * probe_seconds)))
        t0 = time.perf_counter()
        oracle.regf(js.x, js.v, js.m, js.r2, order[:size], js.cap, nthreads=threads)
        rate = size / max(time.perf_counter() - t0, 1e-6)
    count = int(min(js.n, max(size, rate * target_seconds)))
    sample = order[:count]
    t0 = time.perf_counter()
    oracle.regf(js.x, js.v, js.m, js.r2, sample, js.cap, nthreads=threads)
    dt = time.perf_counter() - t0
    inter = float(count) * (js.n - 1)
    info = {"value": inter / dt, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"{count} random targets of the same N={js.n} j-set "
                      f"({inter:.3e} interactions, {dt:.1f} s, FP64 C loop, OpenMP {threads} threads)"}
    # one core on a smaller sample (the reference's CPU loop is serial per
    # particle), sized to ~target_seconds / 2
    n1 = int(min(js.n, max(16, rate / threads * target_seconds / 2)))
    t0 = time.perf_counter()
    oracle.regf(js.x, js.v, js.m, js.r2, order[:n1], js.cap, nthreads=1)
    dt1 = time.perf_counter() - t0
    info["single_core"] = {"value": float(n1) * (js.n - 1) / dt1, "cores": 1,
                           "sample": f"{n1} targets, {dt1:.1f} s"}
    checks = None
    if gpu_results:
        from oracle.compare import compare_lists, force_errors
        t = sample[:nsample]
        ra, rj, rn, rl, sa, sj = oracle.regf(js.x, js.v, js.m, js.r2, t, js.cap, scales=True)
        checks = {}
        for name, (acc, adot, nnb, lst) in gpu_results.items():
            err = force_errors(acc[t], adot[t], ra, rj, sa, sj)
            lists = compare_lists(nnb[t], lst[t], rn, rl, js.x, js.r2, t)
            checks[name] = {
                "sample_targets": int(len(t)), "a_rel_max": err["a_rel"]["max"],
                "a_rel_p99": err["a_rel"]["p99"], "jdot_rel_max": err["jdot_rel"]["max"],
                "jdot_rel_p99": err["jdot_rel"]["p99"], "a_norm_max": err["a_norm"]["max"],
                "jdot_norm_max": err["jdot_norm"]["max"],
                "list_hard_mismatches": lists["hard_mismatches"],
                "list_band_pairs_flipped": lists["band_pairs_flipped"]}
    return info, checks


def full_n_errors(res, ref):
    """Per-particle relative errors of a and jdot over every target against the
    FP64 kernel's results, on the rows whose neighbour set (hence regular set)
    is the same; the other rows differ by a pair within the 1e-6 band."""
    import numpy as np
    acc, adot, nnb, lst = res
    racc, radot, rnnb, rlst = ref
    w = lst.shape[1]
    valid = np.arange(w)[None, :] < np.minimum(np.abs(nnb), w)[:, None]
    rows = (nnb == rnnb) & np.all((lst == rlst) | ~valid, axis=1)
    ea = np.linalg.norm(acc - racc, axis=1)[rows] / np.linalg.norm(racc, axis=1)[rows]
    ej = np.linalg.norm(adot - radot, axis=1)[rows] / np.linalg.norm(radot, axis=1)[rows]
    return {"targets": int(rows.sum()), "band_rows": int((~rows).sum()),
            "a_rel_max": float(ea.max()), "a_rel_p99": float(np.percentile(ea, 99)),
            "jdot_rel_max": float(ej.max()), "jdot_rel_p99": float(np.percentile(ej, 99)),
            "jdot_over_1e-6": int((ej > 1e-6).sum()), "a_over_1e-6": int((ea > 1e-6).sum())}


def precision_leg(ctx, targets, inter, name, reps=5):
    """The same step on the same context and targets in another precision:
    HIP-event time per launch over `reps` launches (checked precision waits
    on the host inside each launch), restoring the context's mode after."""
    import numpy as np
    from abyss_amd import regf
    ev = HipEvents()
    ctx.set_precision(prec_code(name))
    ctx.set_targets(targets)
    ctx.launch()
    ctx.sync()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, ctx.stream)
    for _ in range(reps):
        ctx.launch()
    ev.record(e1, ctx.stream)
    ms = ev.elapsed_ms(e0, e1) / reps
    ctx.sync()
    info = {"ms_per_step": ms, "value": inter / (ms * 1e-3),
            "pct_fp32_peak": 100.0 * FLOP_PER_INTERACTION * inter / (ms * 1e-3) / 1e12
                             / FP32_PEAK_TFLOPS,
            "timing": f"HIP events over {reps} launches"}
    if name == "checked":
        info["recomputed_targets"] = ctx.last_recheck()
        info["tolerance"] = 1e-6
    res = ctx.fetch()
    return info, res


def bench_ensemble(args):
    """Batched ensemble vs per-cluster launches (device-resident and legacy).

    Every leg at the library's default precision (checked: every target
    within 1e-6 of FP64) unless named split.  The reference's own kernel
    shape, one target per thread and GridDimY = 64 contiguous j-ranges
    (src_mpi/def.h:77-79, src_mpi/cuda/cuda_kernels.cu:277-279), exists
    here only as split precision (one target per lane), so it is compared
    with the batch in split precision."""
    import numpy as np
    from abyss_amd import regf
    t0 = time.perf_counter()
    from abyss_amd import workloads
    cl = workloads.ensemble_clusters()
    log(f"ensemble workload: {len(cl)} clusters, sum N = {sum(len(c[2]) for c in cl)}, "
        f"built in {time.perf_counter() - t0:.1f} s")
    tuned = False
    if not getattr(args, "no_autotune", False):
        # the plan table, tuned on the workload's own largest cluster
        import tempfile
        big = max(range(len(cl)), key=lambda k: len(cl[k][2]))
        x, v, m, r2, _ = cl[big]
        t0 = time.perf_counter()
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "plans.txt")
            regf.autotune_jset(0, m, x, v, r2, path)
            tuned = {"tuned_on": f"ensemble cluster {big} (Plummer, N = {len(m)})",
                     "seconds": time.perf_counter() - t0, "measured_buckets": tuned_table(path)}
        log(f"autotuned plan table {tuned}")
    inter = float(sum(len(c[4]) * (len(c[2]) - 1) for c in cl))
    cap = 256
    ev = HipEvents()

    def timed_steps(launch, stream, steps):
        """HIP-event time per step over `steps` launches (the checked
        precision waits on the host inside each launch; that wait is part
        of the step)."""
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream)
        for _ in range(steps):
            launch()
        ev.record(e1, stream)
        return ev.elapsed_ms(e0, e1) / steps

    # (1) batched: one persistent graph launch per step
    ens = regf.Ensemble(device=0, cap=cap, max_clusters=len(cl))
    for k, (x, v, m, r2, t) in enumerate(cl):
        ens.upload(k, m, x, v, r2)
        ens.set_targets(k, t)
    cluster_plans = {}
    for k in range(len(cl)):
        jl, tile = ens.plan(k)
        key = f"jl{jl}/tile{tile}"
        cluster_plans[key] = cluster_plans.get(key, 0) + 1
    log(f"cluster plans (lanes per target / j-tile): {cluster_plans}")
    ms_batched = {}
    for prec, name in ((regf.PREC_CHECKED, "checked"), (regf.PREC_SPLIT, "split")):
        ens.set_precision(prec)
        for _ in range(args.warmup):
            ens.launch()
        ens.sync()
        ms_batched[name] = timed_steps(ens.launch, ens.stream, args.steps)
        ens.sync()
        if name == "checked":
            recomputed = sum(ens.last_recheck(k) for k in range(len(cl)))
        log(f"batched ({name}): {ms_batched[name]:.3f} ms/step")
    # results into preallocated host buffers, as the reference caller keeps
    # its AccRegReceive / ACListReceive arrays (both paths below use them)
    ens.set_precision(regf.PREC_CHECKED)
    bufs = [(np.zeros((len(t), 3)), np.zeros((len(t), 3)), np.zeros(len(t), np.int32),
             np.zeros((len(t), cap), np.int32)) for (_, _, _, _, t) in cl]
    ens.launch()
    ens.sync()
    for k in range(len(cl)):           # warm-up: sizes each cluster's pinned staging
        ens.fetch(k, out=bufs[k])
    w0 = time.perf_counter()
    ens.launch()
    ens.sync()
    res_b = [ens.fetch(k, out=bufs[k]) for k in range(len(cl))]
    ms_batched_fetch = (time.perf_counter() - w0) * 1e3
    log(f"batched (checked) incl. fetch: {ms_batched_fetch:.1f} ms")
    res_b = [tuple(a.copy() for a in r) for r in res_b]   # the buffers are reused below
    ens.close()
    # (2) per-cluster device-resident launches with the library's own plans
    # (lanes per target, j-splits), one after another
    ctxs = []
    for (x, v, m, r2, t) in cl:
        c = regf.Context(device=0, cap=cap)
        c.upload(m, x, v, r2)
        c.set_targets(t)
        ctxs.append(c)
    ms_planned = {}
    for prec, name in ((regf.PREC_CHECKED, "checked"), (regf.PREC_SPLIT, "split")):
        for c in ctxs:
            c.set_precision(prec)
            c.launch()
            c.sync()
        times = []
        for _ in range(max(1, args.steps // 2)):
            w0 = time.perf_counter()
            for c in ctxs:
                c.launch()
                c.sync()
            times.append((time.perf_counter() - w0) * 1e3)
        ms_planned[name] = float(np.median(times))
        log(f"per-cluster planned launches ({name}): {ms_planned[name]:.3f} ms/step")
        if name == "checked":
            # equal accuracy: both meet 1e-6 per target; the batch splits j
            # by work-item size and the planned launch by its own rule, so
            # values differ at rounding level, counts and lists not at all
            lists_equal, max_rel = 0, 0.0
            for c, rb in zip(ctxs, res_b):
                ra = c.fetch()
                w = np.arange(cap)[None, :] < np.minimum(np.abs(ra[2]), cap)[:, None]
                lists_equal += int(np.array_equal(ra[2], rb[2]) and
                                   bool(np.all((ra[3] == rb[3]) | ~w)))
                for xx, yy in zip(ra[:2], rb[:2]):
                    n = np.linalg.norm(xx, axis=1)
                    d = np.linalg.norm(xx - yy, axis=1)
                    max_rel = max(max_rel, float(np.max(d / np.where(n > 0, n, 1.0))))
    # (2b) the reference kernel's shape: one target per lane, 64 contiguous
    # j-ranges per cluster (GridDimY = 64; fewer where N / 64 is under one
    # 256-record tile), split precision
    for c, (_, _, _, _, t) in zip(ctxs, cl):
        c.set_precision(regf.PREC_SPLIT)
        c.set_tuning(1, 64, 2, 1)
        c.set_targets(t)
        c.launch()
        c.sync()
    times = []
    for _ in range(max(1, args.steps // 2)):
        w0 = time.perf_counter()
        for c in ctxs:
            c.launch()
            c.sync()
        times.append((time.perf_counter() - w0) * 1e3)
    ms_refshape = float(np.median(times))
    log(f"per-cluster launches in the reference's shape (split): {ms_refshape:.1f} ms/step")
    for c in ctxs:
        c.close()
    # (3) per-cluster legacy ABI calls (upload + query + results on the host)
    regf.set_num_neighbor_max(cap)
    regf.OpenDevice(0)
    for k, (x, v, m, r2, t) in enumerate(cl):   # warm-up pass (staging sizes)
        regf.SendToDevice(m, x, v, r2)
        regf.CalculateAccelerationOnDevice(t, out=bufs[k])
    w0 = time.perf_counter()
    for k, (x, v, m, r2, t) in enumerate(cl):
        regf.SendToDevice(m, x, v, r2)
        regf.CalculateAccelerationOnDevice(t, out=bufs[k])
    ms_legacy = (time.perf_counter() - w0) * 1e3
    regf.CloseDevice()
    regf.set_num_neighbor_max(1000)
    log(f"per-cluster legacy calls (checked): {ms_legacy:.1f} ms/step")
    return {
        "workload": "ensemble_256_plummer_n1024-65536_active0.01-1",
        "clusters": len(cl), "interactions_per_step": inter, "cap": cap,
        "cluster_plans": cluster_plans,
        "autotuned": tuned,
        "batched_ms_per_step": ms_batched["checked"],
        "batched_split_ms_per_step": ms_batched["split"],
        "interactions_per_s": inter / (ms_batched["checked"] * 1e-3),
        "interactions_per_s_split": inter / (ms_batched["split"] * 1e-3),
        "checked_precision_recomputed_targets": recomputed,
        "per_cluster_planned_launches_ms": ms_planned["checked"],
        "per_cluster_planned_launches_split_ms": ms_planned["split"],
        "per_cluster_reference_shape_split_ms": ms_refshape,
        "per_cluster_legacy_calls_ms": ms_legacy,
        "batched_with_fetch_ms": ms_batched_fetch,
        "speedup_vs_planned_launches": ms_planned["checked"] / ms_batched["checked"],
        "speedup_vs_planned_launches_split": ms_planned["split"] / ms_batched["split"],
        "speedup_vs_reference_shape_split": ms_refshape / ms_batched["split"],
        "speedup_vs_legacy_calls": ms_legacy / ms_batched_fetch,
        "clusters_with_lists_equal_to_per_cluster_launches": lists_equal,
        "max_rel_diff_vs_per_cluster_launches": max_rel,
    }


ADMIT_PAIRS = 3.0e11


def serve_requests(sizes, rate, window_s, seed=9):
    """Arrival times of a Poisson stream at `rate` req/s over `window_s`
    seconds, and its requests: (cluster, targets, N), one regular step of a
    random cluster with an active block log-uniform in [1, N]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * window_s * 1.5) + 16))
    t_arr = t_arr[t_arr < window_s]
    reqs = []
    for _ in t_arr:
        c = int(rng.integers(len(sizes)))
        n = sizes[c]
        mblk = int(min(n, max(1, np.exp(rng.uniform(0.0, np.log(n + 1))))))
        reqs.append((c, rng.choice(n, mblk, replace=False).astype(np.int32), n))
    return t_arr, reqs


def batched_same_mix(cl, reqs, cap):
    """The requests of a served stream as batched-ensemble launches, with no
    server: each launch takes the oldest pending request of every cluster
    that has one (a cluster holds one target set per launch).  Returns the
    interactions/s over the wall time of all launches (targets set, launch,
    wait), checked precision."""
    from abyss_amd import regf
    ens = regf.Ensemble(device=0, cap=cap, max_clusters=len(cl))
    for k, (x, v, m, r2) in enumerate(cl):
        ens.upload(k, m, x, v, r2)
    pending = [[] for _ in cl]
    for (c, t, n) in reqs:
        pending[c].append(t)
    inter = 0.0
    launches = 0
    # warm-up launch of one request (not timed)
    first = next(k for k in range(len(cl)) if pending[k])
    ens.set_targets(first, pending[first][0])
    ens.launch()
    ens.sync()
    t0 = time.perf_counter()
    while any(pending):
        for k in range(len(cl)):
            if pending[k]:
                t = pending[k].pop(0)
                ens.set_targets(k, t)
                inter += float(len(t)) * (len(cl[k][2]) - 1)
            else:
                ens.set_targets(k, [])
        ens.launch()
        ens.sync()
        launches += 1
    wall = time.perf_counter() - t0
    ens.close()
    return {"interactions_per_s": inter / wall, "launches": launches, "wall_s": wall,
            "requests": len(reqs)}


def serve_stream(srv, ids, sizes, rate, window_s, seed=9, check_every=0):
    """Poisson request stream at `rate` req/s for `window_s` seconds (see
    serve_requests).  With check_every = k, every k-th admitted request's
    results are copied back and kept.  Returns (latencies_ms, interactions,
    wall_s, rejected, kept) over the admitted requests, kept = [(cluster
    index, targets, results)]."""
    import threading
    import numpy as np
    t_arr, reqs = serve_requests(sizes, rate, window_s, seed)
    tickets = [None] * len(reqs)
    lat = [None] * len(reqs)
    kept = []

    def waiter():
        q = 0
        while q < len(reqs):
            while tickets[q] is None:
                time.sleep(0.0002)
            tk, nt = tickets[q]
            if tk is not None:   # None: turned away by admission
                if check_every and q % check_every == 0:
                    *res, lat[q] = srv.wait(tk, nt)
                    kept.append((reqs[q][0], reqs[q][1], res))
                else:
                    lat[q] = srv.wait(tk, nt, copy=False)
            q += 1

    wt = threading.Thread(target=waiter)
    wt.start()
    t0 = time.perf_counter()
    for q, (c, t, n) in enumerate(reqs):
        delay = t_arr[q] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets[q] = srv.submit(ids[c], t)
    wt.join()
    wall = time.perf_counter() - t0
    ok = [q for q in range(len(reqs)) if tickets[q][0] is not None]
    inter = float(sum(len(reqs[q][1]) * (reqs[q][2] - 1) for q in ok))
    return np.array([lat[q] for q in ok]), inter, wall, len(reqs) - len(ok), kept


def check_served(cl, kept, nsample=4):
    """a and jdot of `nsample` targets of each kept request against the
    oracle (relative error, rows whose neighbour set equals the oracle's),
    and neighbour counts / lists on the same targets."""
    import numpy as np
    import oracle
    worst_a = worst_j = 0.0
    list_mismatch = 0
    for k, (c, t, (acc, adot, nnb, lst)) in enumerate(kept):
        x, v, m, r2 = cl[c]
        pick = np.random.default_rng(k).permutation(len(t))[:nsample]
        ra, rj, rn, rl = oracle.regf(x, v, m, r2, t[pick], lst.shape[1])
        w = lst.shape[1]
        valid = np.arange(w)[None, :] < np.minimum(np.abs(nnb[pick]), w)[:, None]
        rows = (nnb[pick] == rn) & np.all((lst[pick] == rl) | ~valid, axis=1)
        list_mismatch += int((~rows).sum())
        if rows.any():
            ea = np.linalg.norm(acc[pick] - ra, axis=1) / np.linalg.norm(ra, axis=1)
            ej = np.linalg.norm(adot[pick] - rj, axis=1) / np.linalg.norm(rj, axis=1)
            worst_a = max(worst_a, float(ea[rows].max()))
            worst_j = max(worst_j, float(ej[rows].max()))
    return {"requests_checked": len(kept), "targets_per_request": nsample,
            "a_rel_max": worst_a, "jdot_rel_max": worst_j,
            "rows_differing_from_oracle": list_mismatch}


def tuned_table(path):
    """The measured buckets of a table written by abyss_autotune (the target
    counts 1, 4, 16, ... fall in even buckets): bucket -> "lanes/tile/waves/
    workgroups per CU"."""
    rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
    return {f"2^{r[0]}": f"jl{r[1]}/tile{r[4]}/w{r[5]}/wg{r[3]}" for r in rows
            if int(r[0]) % 2 == 0 and int(r[0]) <= 18}


def bench_tuning(args):
    """The plan table tuned on a uniform cube (N = 65536, what the default
    run installs) against one tuned on each workload's own cluster, at the
    library's default precision: per case, the device-resident launch time
    under each table (HIP events, mean over repeated launches).  Cases: the
    Plummer + Kroupa N = 65536 active blocks and the King W0 = 6 + binaries
    N = 262144 blocks and full-N step."""
    import tempfile
    import numpy as np
    from abyss_amd import regf, workloads
    ev = HipEvents()
    cases = [("plummer_kroupa65536", workloads.plummer_kroupa(65536),
              (1, 16, 256, 4096, 65536)),
             ("king_w6_binaries262144", workloads.king_binaries(262144),
              (1, 256, 16384, 262144))]

    def timed(js, sizes):
        out = {}
        ctx = regf.Context(device=0, cap=js.cap)
        ctx.upload(js.m, js.x, js.v, js.r2)
        rng = np.random.default_rng(11)
        for m in sizes:
            t = np.sort(rng.permutation(js.n)[:m]).astype(np.int32)
            reps = int(min(400, max(5, 0.3 / max(m * (js.n - 1) / 1.0e13, 1e-5))))
            ctx.set_targets(t)
            ctx.launch()
            ctx.sync()
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, ctx.stream)
            for _ in range(reps):
                ctx.launch()
            ev.record(e1, ctx.stream)
            p = ctx.plan()
            out[m] = {"us": ev.elapsed_ms(e0, e1) / reps * 1e3,
                      "plan": f"jl{p['lanes_per_target']}/tile{p['tile']}/"
                              f"w{p['waves_per_block']}/S{p['j_splits']}"}
            ctx.sync()
        ctx.close()
        return out

    res = {}
    with tempfile.TemporaryDirectory() as td:
        cube = os.path.join(td, "cube.txt")
        t0 = time.perf_counter()
        regf.autotune(0, 65536, cube)
        log(f"cube table in {time.perf_counter() - t0:.1f} s: {tuned_table(cube)}")
        res["cube_table"] = tuned_table(cube)
        for name, js, sizes in cases:
            regf.autotune_load(0, cube)
            a = timed(js, sizes)
            own = os.path.join(td, name + ".txt")
            t0 = time.perf_counter()
            regf.autotune_jset(0, js.m, js.x, js.v, js.r2, own)
            secs = time.perf_counter() - t0
            b = timed(js, sizes)
            rows = [{"m": m, "cube_us": a[m]["us"], "cube_plan": a[m]["plan"],
                     "own_us": b[m]["us"], "own_plan": b[m]["plan"],
                     "cube_over_own": a[m]["us"] / b[m]["us"]} for m in sizes]
            for r in rows:
                log(f"{name} m={r['m']}: cube {r['cube_us']:.1f} us ({r['cube_plan']}) "
                    f"own {r['own_us']:.1f} us ({r['own_plan']}) ratio {r['cube_over_own']:.3f}")
            res[name] = {"own_table": tuned_table(own), "own_tuning_s": secs, "rows": rows}
    regf.autotune_clear(0)
    return res


def bench_blocks(args, sizes=(1, 4, 16, 64, 256, 1024, 4096, 16384, 65536)):
    """Active-block sweep on Plummer + Kroupa N = 65536 (cap 256, ~100
    neighbours): per block size m, the device-resident launch (HIP events on
    the context's stream, inputs in HBM) and the synchronous query (targets
    to the device, launch, results into the caller's preallocated host
    arrays; wall clock, median), with interactions/s, the
    FP32 VALU fraction and the HBM fraction of the algorithmic bytes
    68 B * N (j-set) + 64 B * m (targets) + (52 + 4 n_nb) B * m (results)."""
    import numpy as np
    from abyss_amd import regf, workloads
    js = workloads.plummer_kroupa(65536)
    n = js.n
    prec = prec_code(args.precision)
    ctx = regf.Context(device=0, cap=js.cap, precision=prec)
    ctx.upload(js.m, js.x, js.v, js.r2)
    regf.set_num_neighbor_max(js.cap)
    regf.set_precision(prec)
    regf.OpenDevice(0)
    ev = HipEvents()
    rng = np.random.default_rng(7)
    rows = []
    blocks = []
    for m in sizes:
        t = np.sort(rng.permutation(n)[:m]).astype(np.int32)
        blocks.append(t)
        inter = float(m) * (n - 1)
        reps = int(min(2000, max(10, 0.2 / max(inter / 1.2e12, 1e-5))))
        ctx.set_targets(t)
        ctx.launch()
        ctx.sync()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, ctx.stream)
        for _ in range(reps):
            ctx.launch()
        ev.record(e1, ctx.stream)
        ms = ev.elapsed_ms(e0, e1) / reps
        ctx.sync()
        nnb = ctx.fetch()[2]
        out = (np.zeros((m, 3)), np.zeros((m, 3)), np.zeros(m, np.int32),
               np.zeros((m, js.cap), np.int32))     # the caller's result arrays
        ctx.query(t, out=out)
        walls = []
        for _ in range(min(reps, 200)):
            w0 = time.perf_counter()
            ctx.query(t, out=out)
            walls.append((time.perf_counter() - w0) * 1e3)
        q_ms = float(np.median(walls))
        byts = 68.0 * n + 64.0 * m + (52.0 + 4.0 * float(np.abs(nnb).mean())) * m
        # the reference's calling pattern: the whole j-set uploaded, then the
        # block queried, through the legacy entry points
        regf.SendToDevice(js.m, js.x, js.v, js.r2)
        regf.CalculateAccelerationOnDevice(t, out=out)
        lwalls = []
        for _ in range(min(reps, 50)):
            w0 = time.perf_counter()
            regf.SendToDevice(js.m, js.x, js.v, js.r2)
            regf.CalculateAccelerationOnDevice(t, out=out)
            lwalls.append((time.perf_counter() - w0) * 1e3)
        l_ms = float(np.median(lwalls))
        row = {"m": m, "plan": ctx.plan(), "launch_us": ms * 1e3, "query_us": q_ms * 1e3,
               "legacy_upload_and_call_us": l_ms * 1e3,
               "interactions_per_s": inter / (ms * 1e-3),
               "valu_frac": FLOP_PER_INTERACTION * inter / (ms * 1e-3) / (FP32_PEAK_TFLOPS * 1e12),
               "hbm_bytes": byts, "hbm_frac": byts / (ms * 1e-3) / HBM_PEAK_BPS}
        log(f"m={m:6d}: launch {row['launch_us']:9.1f} us, query {row['query_us']:9.1f} us, "
            f"legacy upload+call {row['legacy_upload_and_call_us']:9.1f} us, "
            f"{row['interactions_per_s']:.3e} int/s, VALU {100 * row['valu_frac']:5.1f} %, "
            f"HBM {100 * row['hbm_frac']:5.2f} %, plan {row['plan']}")
        rows.append(row)
    ctx.close()
    regf.CloseDevice()
    regf.set_num_neighbor_max(1000)
    regf.set_precision(regf.PREC_DEFAULT)
    from_c = legacy_from_c(js, blocks, prec)
    for row, c in zip(rows, from_c or []):
        row["legacy_from_c_us"] = c
    if from_c:
        log("legacy upload+call from C: " +
            ", ".join(f"m={r['m']} {r['legacy_from_c_us']['total_us']:.1f} us" for r in rows))
    return {"workload": "plummer_kroupa_n65536_cap256", "precision": args.precision,
            "rows": rows}


def config1_hermite(t_end=1.0, eta=0.01):
    """The plain-method workload's dynamical check: the equal-mass Plummer
    sphere N = 1024 (seed 1, ~48 neighbours) integrated for one N-body time
    by tests/harness/ac_hermite_gpu, a block-timestep Hermite integrator
    whose regular force and neighbour lists come from the six legacy entry
    points (checked precision), as ABYSS's caller drives them.  Returns its
    JSON line (energy error dE/E, steps, wall time) or None when the harness
    is not built."""
    import struct
    import subprocess
    import tempfile
    from abyss_amd import workloads
    exe = os.path.join(ROOT, "tests", "harness", "ac_hermite_gpu")
    if not os.path.exists(exe):
        return None
    js = workloads.plummer_equal(1024)
    with tempfile.TemporaryDirectory() as td:
        ic = os.path.join(td, "ic.bin")
        with open(ic, "wb") as f:
            f.write(struct.pack("<i", js.n))
            for a in (js.x, js.v, js.m, js.r2):
                f.write(a.astype("<f8").tobytes())
        p = subprocess.run([exe, "gpu", ic, str(t_end), str(eta)], capture_output=True,
                           text=True, timeout=300)
    if p.returncode != 0:
        log(f"Hermite harness failed ({p.returncode}): {p.stderr[-300:]}")
        return None
    out = json.loads(p.stdout.strip().splitlines()[-1])
    out["workload"] = "plummer_equal_n1024_seed1, 1 N-body time"
    return out


def legacy_from_c(js, blocks, prec):
    """The legacy calling pattern (SendToDevice of the whole j-set, then
    CalculateAccelerationOnDevice of the block) timed from C, as ABYSS calls
    it, by tests/harness/legacy_timing on the same j-set and blocks: per
    block {send_us, call_us, total_us} (medians), or None when the harness
    is not built."""
    import subprocess
    import tempfile
    import numpy as np
    exe = os.path.join(ROOT, "tests", "harness", "legacy_timing")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "workload.bin")
        with open(path, "wb") as f:
            np.array([js.n], np.int32).tofile(f)
            for a in (js.m, js.x, js.v, js.r2):
                np.ascontiguousarray(a, np.float64).tofile(f)
            np.array([len(blocks)], np.int32).tofile(f)
            for t in blocks:
                np.array([len(t)], np.int32).tofile(f)
                np.ascontiguousarray(t, np.int32).tofile(f)
        p = subprocess.run([exe, path, str(js.cap), str(prec), "100"], capture_output=True,
                           text=True, timeout=120)
    if p.returncode != 0:
        log(f"legacy timing harness failed ({p.returncode}): {p.stderr[-300:]}")
        return None
    out = [json.loads(ln) for ln in p.stdout.splitlines() if ln.startswith("{")]
    return out if len(out) == len(blocks) else None


def bench_neighbour(args, reps=5):
    """Irregular force over resident lists and the fused regular step
    (neighbour-set correction + corrector) on the King N=262144 workload,
    all N particles per call; wall-clock per synchronous call."""
    import numpy as np
    from abyss_amd import regf, workloads
    js = workloads.king_binaries(args.n)
    t = np.arange(js.n, dtype=np.int32)
    rng = np.random.default_rng(3)
    dt = np.full(js.n, 1.0 / 1024)
    a_old = rng.normal(size=(js.n, 6))
    out = {"workload": f"{js.name}_fullN", "n": js.n}
    with regf.Context(device=0, cap=js.cap) as ctx:
        ctx.keep_lists(True)
        ctx.upload(js.m, js.x, js.v, js.r2)
        _, _, nnb, _ = ctx.query(t)
        entries = int(np.minimum(np.abs(nnb), js.cap).sum())

        def wall(f):
            f()
            t0 = time.perf_counter()
            for _ in range(reps):
                f()
            return (time.perf_counter() - t0) / reps * 1e3

        ms_irr = wall(lambda: ctx.irregular(t))
        ms_query = wall(lambda: ctx.query(t))
        ms_step = wall(lambda: ctx.regular_step(t, dt, a_old))
        ms_radius = wall(lambda: ctx.update_radius(100, 1000))
        # j prediction: host predictor + pack + H2D (upload) vs the device
        # predictor over the resident state
        a = rng.normal(size=(js.n, 3)) * 0.1
        ad = rng.normal(size=(js.n, 3)) * 0.1
        t0 = np.zeros(js.n)
        ctx.upload_state(js.m, js.x, js.v, a, ad, t0, js.r2)

        def host_predict_upload():
            h = 1.0 / 1024
            xp = ((ad * h / 3 + a) * h / 2 + js.v) * h + js.x
            vp = (ad * h / 2 + a) * h + js.v
            ctx.upload(js.m, xp, vp, js.r2)
            ctx.sync()

        def device_predict():
            ctx.predict(1.0 / 1024)
            ctx.sync()

        ms_host_pred = wall(host_predict_upload)
        ctx.upload_state(js.m, js.x, js.v, a, ad, t0, js.r2)
        ms_dev_pred = wall(device_predict)
        # start-up derivatives over all N (one call)
        t1 = time.perf_counter()
        ctx.derivs23(t)
        ms_d23 = (time.perf_counter() - t1) * 1e3
    out["irregular"] = {"targets": js.n, "list_entries": entries, "ms_per_call": ms_irr,
                        "entries_per_s": entries / (ms_irr * 1e-3),
                        "includes": "target H2D, kernel, result D2H"}
    out["regular_step"] = {"targets": js.n, "ms_per_call": ms_step, "query_ms_per_call": ms_query,
                           "correction_overhead_ms": ms_step - ms_query}
    out["update_radius"] = {"targets": js.n, "ms_per_call": ms_radius}
    out["prediction"] = {"host_predict_pack_upload_ms": ms_host_pred,
                         "device_predict_ms": ms_dev_pred}
    out["derivs23"] = {"targets": js.n, "pairs": float(js.n) * (js.n - 1), "ms": ms_d23,
                       "pairs_per_s": float(js.n) * (js.n - 1) / (ms_d23 * 1e-3)}
    log(f"neighbour legs {out}")
    return out


def bench_serve(args, rate=2000.0, window_s=1.0):
    """Served ensemble (64 clusters, N log-uniform in [4096, 1048576]) at the
    library's default precision (checked): Poisson streams at the nominal
    2000 req/s with admission, and at 100 and 300 req/s without; p50/p99
    latency, interactions/s, the GPU-busy fraction, and an oracle check of a
    sample of the served results."""
    import numpy as np
    from abyss_amd import regf
    t0 = time.perf_counter()
    from abyss_amd import workloads
    cl = workloads.served_clusters()
    sizes = [len(c[2]) for c in cl]
    log(f"served workload: {len(cl)} clusters, sum N = {sum(sizes)}, built in "
        f"{time.perf_counter() - t0:.1f} s")
    cap = 128
    srv = regf.Server(device=0, cap=cap, max_clusters=len(cl), hbm_budget=160 << 30)
    ids = [srv.add_cluster(m, x, v, r2) for (x, v, m, r2) in cl]
    # warm-up: one small request per cluster
    for c in ids:
        tk, nt = srv.submit(c, np.arange(8, dtype=np.int32))
        srv.wait(tk, nt, copy=False)
    out = {"precision": "checked"}
    # nominal_admission: requests that would take the outstanding work past
    # ADMIT_PAIRS (~0.25 s of the chip's work) are turned away; 100 and 300
    # req/s: every request queues
    legs = (("nominal_admission", rate, window_s, ADMIT_PAIRS, 0),
            ("rate100", args.serve_rate, 4.0, 0.0, 10),
            ("rate300", 3 * args.serve_rate, 3.0, 0.0, 0))
    kept_all = []
    for name, r, w, admit, check_every in legs:
        srv.set_admission(admit)
        srv.reset_stats()
        lat, inter, wall, rejected, kept = serve_stream(srv, ids, sizes, r, w,
                                                        check_every=check_every)
        st = srv.stats()
        kept_all += kept
        out[name] = {"rate_req_s": r, "window_s": w, "admission_pairs": admit or None,
                     "requests": int(len(lat)) + rejected, "admitted": int(len(lat)),
                     "rejected": rejected, "max_queued": st["max_pending"],
                     "peak_outstanding_pairs": st["peak_outstanding_pairs"],
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "interactions_per_s": inter / wall, "wall_s": wall, "batches": st["batches"],
                     "mean_batch_size": st["completed"] / max(1, st["batches"]),
                     "gpu_busy_frac": st["gpu_busy_frac"]}
        log(f"serve {name}: {out[name]}")
    srv.close()
    # the rate300 leg's requests as batched launches without the server: the
    # batched rate of the served mix itself
    _, reqs300 = serve_requests(sizes, 3 * args.serve_rate, 3.0)
    out["batched_same_mix"] = batched_same_mix(cl, reqs300, cap)
    out["rate300_over_batched_same_mix"] = (out["rate300"]["interactions_per_s"]
                                            / out["batched_same_mix"]["interactions_per_s"])
    log(f"serve batched same mix: {out['batched_same_mix']} "
        f"(rate300 / batched {out['rate300_over_batched_same_mix']:.3f})")
    out["oracle_sample"] = check_served(cl, kept_all)
    log(f"serve oracle sample: {out['oracle_sample']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--precision", choices=PREC_NAMES, default="checked",
                    help="precision of the timed step (default: checked, the library "
                         "default, every target within 1e-6 of FP64)")
    ap.add_argument("--ipl", type=int, default=0)
    ap.add_argument("--splits", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble and served-ensemble legs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the built-in plan rules instead of tuning first")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy legs")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes that measure HBM traffic")
    ap.add_argument("--no-blocks", action="store_true",
                    help="skip the compact active-block sweep (N = 65536)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration)")
    ap.add_argument("--workload", choices=["king262144", "ensemble", "serve", "neighbour",
                                           "blocks", "tuning"],
                    default="king262144")
    ap.add_argument("--serve-rate", type=float, default=100.0,
                    help="sustained request rate of the served-ensemble leg (req/s)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.workload == "ensemble":
        print(json.dumps({"metric": "batched ensemble regular-force pair-interactions/s",
                          "batched_ensemble": bench_ensemble(args)}), flush=True)
        return
    if args.workload == "blocks":
        print(json.dumps({"metric": "regular-force active-block sweep (latency and rates)",
                          "blocks": bench_blocks(args)}), flush=True)
        return
    if args.workload == "tuning":
        print(json.dumps({"metric": "plan tables: cube-tuned vs tuned on the workload",
                          "tuning": bench_tuning(args)}), flush=True)
        return
    if args.workload == "neighbour":
        print(json.dumps({"metric": "irregular force and fused regular step (wall-clock)",
                          "neighbour": bench_neighbour(args)}), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps({"metric": "served ensemble regular-step latency",
                          "served_ensemble": bench_serve(args)}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the regular-force path runs on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # HBM traffic of the force kernel: counter passes in child processes,
    # before this process opens the device
    traffic = traffic_why = None
    if not args.no_traffic and args.n == 262144:
        t0 = time.perf_counter()
        traffic, traffic_why = live_traffic(args.precision)
        log(f"traffic {traffic or traffic_why} ({time.perf_counter() - t0:.1f} s)")

    from abyss_amd import regf, workloads

    t0 = time.perf_counter()
    js = workloads.king_binaries(args.n)
    log(f"workload {js.name}: built in {time.perf_counter() - t0:.1f} s")
    import numpy as np
    targets = np.arange(js.n, dtype=np.int32)
    ctx = regf.Context(device=0, cap=js.cap, precision=prec_code(args.precision))
    ctx.set_tuning(args.ipl, args.splits)
    ctx.upload(js.m, js.x, js.v, js.r2)
    ctx.set_targets(targets)
    plan = ctx.plan()
    log(f"plan {plan}")
    for _ in range(args.warmup):
        ctx.launch()
    ctx.sync()
    log("warmup done")
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    stream = ctx.stream
    ctx.kernel_timing(True)   # events around the force kernel of every launch
    w0 = time.perf_counter()
    ev.record(e0, stream)
    for _ in range(args.steps):
        ctx.launch()
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    ctx.sync()
    wall = time.perf_counter() - w0
    kt = ctx.read_kernel_timing()
    ctx.kernel_timing(False)
    recomputed = ctx.last_recheck() if args.precision == "checked" else None
    ms_per_step = ms / args.steps
    inter = float(len(targets)) * (js.n - 1)
    value = inter / (ms_per_step * 1e-3)
    flop = FLOP_PER_INTERACTION * inter
    step_tflops = flop / (ms_per_step * 1e-3) / 1e12
    kernel_ms = kt["force_ms"] / max(1, kt["launches"])
    achieved = flop / (kernel_ms * 1e-3) / 1e12
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall:.3f} s), "
        f"{value:.4e} interactions/s, {step_tflops:.2f} TFLOP/s; force kernel "
        f"{kernel_ms:.3f} ms ({achieved:.2f} TFLOP/s); recomputed {recomputed}")

    acc_info = None
    legs = {}
    gpu_results = {}
    if not args.no_check:
        gpu_results[args.precision] = ctx.fetch()
        # the other modes on the same targets, for their time and accuracy
        for other in ("split", "checked"):
            if other != args.precision:
                legs[other], gpu_results[other] = precision_leg(ctx, targets, inter, other)
                log(f"{other} precision {legs[other]}")
        ctx.set_precision(regf.PREC_FP64)
        f64 = ctx.query(targets)
        acc_info = {"full_n_vs_fp64_kernel": {k: full_n_errors(r, f64)
                                              for k, r in gpu_results.items()}}
        log(f"accuracy {acc_info}")
    ctx.close()

    cpu = None
    if not args.no_cpu:
        cpu, checks = cpu_baseline(js, args.cpu_seconds, gpu_results)
        log(f"cpu baseline {cpu}")
        if checks:
            acc_info["sample_vs_oracle"] = checks
            log(f"oracle sample {checks}")
    blocks = None
    if not args.no_blocks:
        blocks = bench_blocks(args, sizes=(1, 64, 4096, 65536))
    hermite = None
    if not args.no_check:
        # the dynamical check of the plain-method workload (an accuracy leg)
        hermite = config1_hermite()
        log(f"Hermite run, Plummer N = 1024: {hermite}")
    ens_info = serve_info = None
    if not args.no_ensemble:
        # the ensemble legs mix every block size: they plan from a table
        # tuned on the ensemble's own largest cluster (the full-N step above
        # uses the built-in rules, which were set from sweeps of this
        # workload); the served leg keeps that table
        ens_args = argparse.Namespace(steps=4, warmup=1, no_autotune=args.no_autotune)
        ens_info = bench_ensemble(ens_args)
        log(f"batched ensemble {ens_info}")
        serve_info = bench_serve(args)

    arith = {"checked": "f64 i-j differences, f32 r^2 / rsqrt / terms, f32 runs of 64 pairs "
                        "folded into f64 sums; targets whose rounding may exceed 1e-6 "
                        "recomputed with f64 terms",
             "split": "f64 i-j differences, f32 r^2 / rsqrt / terms, f32 runs of 64 pairs "
                      "folded into f64 sums",
             "fp64": "f64 throughout"}[args.precision]
    # the dominant kernel as the trace names it (template arguments from the plan)
    if args.precision == "fp64":
        kname = f"regf_kernel<{plan['i_per_lane']}, 2, true>"
    else:
        kname = (f"regf_kernel_jl<{plan['lanes_per_target']}, 0, "
                 f"{'true' if args.precision == 'checked' else 'false'}, {plan['tile']}, "
                 f"{plan['waves_per_block']}>")
    line = {
        "metric": "regular-force pair-interactions/s",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{js.name}_fullN", "n": js.n, "active": int(len(targets)),
                   "neighbours_target": js.meta["k_nb"], "cap": js.cap,
                   "precision": args.precision, "arithmetic": arith, "plan": plan,
                   "recomputed_targets": recomputed},
        "pct_fp32_peak": 100.0 * step_tflops / FP32_PEAK_TFLOPS,
        "roofline": {"bound": "valu", "kernel": kname,
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "kernel_ms": kernel_ms,
                     "kernel_timing": f"HIP events around the force kernel on the context "
                                      f"stream, {kt['launches']} timed launches",
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "interactions_per_launch": inter},
        "traffic_detail": traffic if traffic else {"unmeasured": traffic_why},
        "cpu_baseline": cpu,
        "accuracy": acc_info,
        "other_precisions": legs or None,
        "active_blocks": blocks,
        "config1_hermite": hermite,
        "batched_ensemble": ens_info,
        "served_ensemble": serve_info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
