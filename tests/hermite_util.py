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
    if not os.path.exists(HARNESS):
        from tests.harness import build_harness
        build_harness.build()
    r = subprocess.run([HARNESS, provider, ic_path, str(t_end), str(eta)], capture_output=True,
                       text=True, timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return json.loads(r.stdout.strip().splitlines()[-1])
