"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/abyss_regf.h, and validates before touching the device."""
import os
import ctypes
import re
import subprocess

import numpy as np
import pytest

from abyss_amd import regf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "abyss_regf.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", "", src, flags=re.M)   # preprocessor lines
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def test_header_lists_match_python_binding():
    assert sorted(declared_functions()) == sorted(regf.EXPORTS)


LIB_F32 = regf.LIB_PATH.replace("libabyss_regf.so", "libabyss_regf_f32.so")


@pytest.mark.parametrize("path", [regf.LIB_PATH, LIB_F32])
def test_library_exports_every_declared_symbol(path):
    """Both builds (FP64 and FP32 legacy scalar) export the whole header."""
    assert os.path.exists(path), path
    lib = regf.load() if path == regf.LIB_PATH else ctypes.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in declared_functions():
        assert name in exported, name


def test_header_compiles_as_c_and_cpp(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "abyss_regf.h"\nint main(void){return abyss_device_count()<0;}\n')
    for comp in (["gcc", "-std=c99"], ["g++", "-x", "c++"]):
        subprocess.run(comp + ["-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(c)],
                       check=True)


def test_error_strings():
    assert regf.error_string(regf.OK) == "ok"
    assert "overflow" in regf.error_string(regf.ERR_NEIGHBOR_OVERFLOW)


@pytest.mark.skipif(regf.device_count() > 0, reason="checks the no-device path")
def test_calls_without_device_fail_cleanly():
    with pytest.raises(regf.AbyssError) as e:
        regf.SendToDevice(np.ones(2), np.zeros((2, 3)), np.zeros((2, 3)), np.ones(2))
    assert e.value.code == regf.ERR_NOT_OPEN
    with pytest.raises(regf.AbyssError) as e:
        regf.CalculateAccelerationOnDevice([0])
    assert e.value.code == regf.ERR_NOT_OPEN
    with pytest.raises(regf.AbyssError) as e:
        regf.OpenDevice(0)
    assert e.value.code == regf.ERR_NO_DEVICE
    with pytest.raises(regf.AbyssError) as e:
        regf.Context()
    assert e.value.code == regf.ERR_NO_DEVICE


def test_precision_setters_validate_without_a_context():
    lib = regf.load()
    assert lib.abyss_set_precision(regf.PREC_CHECKED) == regf.OK
    assert lib.abyss_set_precision(regf.PREC_CHECKED + 1) == regf.ERR_INVALID_ARG
    assert lib.abyss_set_precision(-1) == regf.ERR_INVALID_ARG
    assert lib.abyss_set_precision(regf.PREC_SPLIT) == regf.OK
    assert lib.abyss_ctx_set_precision(None, regf.PREC_CHECKED) == regf.ERR_INVALID_ARG
    assert lib.abyss_ctx_set_check_tolerance(None, 1e-6) == regf.ERR_INVALID_ARG
    n = ctypes.c_int(7)
    assert lib.abyss_ctx_last_recheck(None, ctypes.byref(n)) == regf.ERR_INVALID_ARG
    buf = (ctypes.c_float * 2)()
    assert lib.abyss_ctx_fetch_spread(None, buf) == regf.ERR_INVALID_ARG


def test_neighbor_stride_setter_validates():
    assert regf.num_neighbor_max() == 1000
    with pytest.raises(regf.AbyssError):
        regf.set_num_neighbor_max(0)
    regf.set_num_neighbor_max(200)
    assert regf.num_neighbor_max() == 200
    regf.set_num_neighbor_max(1000)


def test_c_caller_links_and_reports_no_device(tmp_path):
    """A plain C caller in the style of ABYSS's host code links against the
    library and gets a clean status (no GPU in this container)."""
    src = tmp_path / "caller.c"
    src.write_text(r'''
#include <stdio.h>
#include "abyss_regf.h"
int main(void) {
    int rank = 0, n = 2, nt = 1, idx[1] = {0}, nnb[1];
    double m[2] = {1, 1}, x[2][3] = {{0,0,0},{1,0,0}}, v[2][3] = {{0}}, r2[2] = {0, 0};
    double acc[1][3], adot[1][3];
    int list[ABYSS_NUM_NEIGHBOR_MAX];
    if (abyss_device_count() > 0) return 0;
    if (abyss_set_legacy_errors(ABYSS_LEGACY_RETURN) != ABYSS_OK) return 4;
    OpenDevice(&rank);
    if (abyss_last_error() != ABYSS_ERR_NO_DEVICE) return 1;
    SendToDevice(&n, m, x, v, r2, NULL);
    if (abyss_last_error() != ABYSS_ERR_NOT_OPEN) return 2;
    CalculateAccelerationOnDevice(&nt, idx, acc, adot, nnb, list);
    if (abyss_last_error() != ABYSS_ERR_NOT_OPEN) return 3;
    printf("ok\n");
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(regf.LIB_PATH)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                    "-labyss_regf", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)


@pytest.mark.parametrize("lib_name,flag", [("abyss_regf", ""), ("abyss_regf_f32",
                                                                "-DABYSS_LEGACY_FLOAT")])
def test_c_caller_stops_on_error_by_default(tmp_path, lib_name, flag):
    """The reference's callers never check a status; by default a failing
    legacy call prints its error and ends the process with exit(1), as the
    reference does (exit on overflow, asserts on a closed device)."""
    src = tmp_path / "caller.c"
    src.write_text(r'''
#include <stdio.h>
#include "abyss_regf.h"
int main(void) {
    int nt = 1, idx[1] = {0}, nnb[1];
    abyss_legacy_real acc[1][3], adot[1][3];
    int list[ABYSS_NUM_NEIGHBOR_MAX];
    printf("before\\n");
    fflush(stdout);
    CalculateAccelerationOnDevice(&nt, idx, acc, adot, nnb, list);   /* not open */
    printf("after\\n");
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(regf.LIB_PATH)
    cmd = ["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir, f"-l{lib_name}",
           f"-Wl,-rpath,{libdir}", "-o", str(exe)]
    if flag:
        cmd.insert(1, flag)
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "before" in r.stdout and "after" not in r.stdout
    assert "device not open" in r.stderr


def test_legacy_error_mode_setter_validates():
    lib = regf.load()
    assert lib.abyss_set_legacy_errors(7) == regf.ERR_INVALID_ARG
    assert lib.abyss_set_legacy_errors(regf.LEGACY_RETURN) == regf.OK


def _table(path, rows):
    with open(path, "w") as f:
        f.write("# bucket jl unroll wg_per_cu [tile waves]\n")
        for r in rows:
            f.write(" ".join(str(v) for v in r) + "\n")


def test_plan_table_files_load_and_validate(tmp_path):
    """abyss_autotune_load parses on the host: six-column tables (lanes per
    target, j-unroll, workgroups per CU, j-tile, waves per workgroup) and
    four-column ones written before the shape columns load; a table with an
    unsupported shape, lanes count or a missing bucket is refused."""
    lib = regf.load()
    shapes = [(256, 4), (128, 4), (64, 4), (128, 2), (64, 2), (64, 1), (1024, 4)]
    good6 = str(tmp_path / "six.txt")
    _table(good6, [(b, (4, 16, 64)[b % 3], 2, 16, *shapes[b % len(shapes)]) for b in range(32)])
    good4 = str(tmp_path / "four.txt")
    _table(good4, [(b, 4, 2, 48) for b in range(32)])
    try:
        assert lib.abyss_autotune_load(0, good6.encode()) == regf.OK
        assert lib.abyss_autotune_load(0, good4.encode()) == regf.OK
        for bad_row, name in (((5, 4, 2, 16, 512, 4), "tile 512"),
                              ((5, 4, 2, 16, 256, 2), "tile above 64 x waves"),
                              ((5, 4, 2, 16, 1024, 2), "one-shot tile with 2 waves"),
                              ((5, 8, 2, 16, 256, 4), "8 lanes per target"),
                              ((5, 4, 3, 16, 256, 4), "unroll 3")):
            path = str(tmp_path / "bad.txt")
            _table(path, [bad_row if b == 5 else (b, 4, 2, 16, 256, 4) for b in range(32)])
            assert lib.abyss_autotune_load(0, path.encode()) == regf.ERR_INVALID_INPUT, name
        short = str(tmp_path / "short.txt")
        _table(short, [(b, 4, 2, 16, 256, 4) for b in range(31)])
        assert lib.abyss_autotune_load(0, short.encode()) == regf.ERR_INVALID_INPUT
        assert lib.abyss_autotune_load(0, str(tmp_path / "missing.txt").encode()) == \
            regf.ERR_INVALID_ARG
        assert lib.abyss_autotune_load(-1, good6.encode()) == regf.ERR_INVALID_ARG
    finally:
        assert lib.abyss_autotune_clear(0) == regf.OK
