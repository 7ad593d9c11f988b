"""Runs the Hermite harness (tests/harness/ac_hermite) on a j-set."""
import json
import os
import struct
import subprocess

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness", "ac_hermite")


def write_ic(path, js):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", js.n))
        for a in (js.x, js.v, js.m, js.r2):
            f.write(a.astype("<f8").tobytes())


def run(provider, ic_path, t_end, eta=0.01, timeout=600):
    return _run(HARNESS, [provider, ic_path, str(t_end), str(eta)], timeout)


AC_FULL = os.path.join(os.path.dirname(HARNESS), "ac_full")


def run_ac(provider, ic_path, t_end, eta_irr, eta_reg, fix, timeout=600):
    """Two-time-scale Ahmad-Cohen harness (tests/harness/ac_full)."""
    return _run(AC_FULL, [provider, ic_path, str(t_end), str(eta_irr), str(eta_reg), str(fix)],
                timeout)


def _run(exe, args, timeout):
    if not os.path.exists(exe):
        from tests.harness import build_harness
        build_harness.build()
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return json.loads(r.stdout.strip().splitlines()[-1])
