"""Host code under the sanitizers, on the CPU: the j-set validation and
multithreaded pack of an upload (abyss_amd/csrc/regf_pack.cpp) and the
oracle's C restatement, built with AddressSanitizer + UndefinedBehavior-
Sanitizer and (the pack) with ThreadSanitizer, driven by
tests/harness/sanitize_check.cpp."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "harness", "sanitize_check.cpp")
PACK = os.path.join(ROOT, "abyss_amd", "csrc", "regf_pack.cpp")
ORACLE = os.path.join(ROOT, "oracle", "regf_oracle.c")

pytestmark = pytest.mark.skipif(not (shutil.which("gcc") and shutil.which("g++")),
                                reason="needs gcc / g++")


def build(tmp_path, name, san, openmp):
    omp = ["-fopenmp"] if openmp else []
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={san}"]
    obj = str(tmp_path / f"{name}_oracle.o")
    subprocess.run(["gcc", "-std=c99", *flags, *omp, "-c", ORACLE, "-o", obj], check=True)
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-pthread", *flags, *omp, DRIVER, PACK, obj, "-lm",
                    "-o", exe], check=True)
    return exe


def run(exe, arg):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1", OMP_NUM_THREADS="2")
    p = subprocess.run([exe, arg], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout[-2000:], p.stderr[-4000:])
    assert p.returncode == 0 and "OK: 0 failure(s)" in p.stdout


def test_pack_and_oracle_under_asan_ubsan(tmp_path):
    run(build(tmp_path, "asan", "address,undefined", openmp=True), "all")


def test_pack_under_tsan(tmp_path):
    run(build(tmp_path, "tsan", "thread", openmp=False), "pack")
