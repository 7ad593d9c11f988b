"""The drop-in recipes compile: ABYSS's own callers, with exactly the edits the
integration recipe lists, pass `g++ -fsyntax-only` against include/abyss_regf.h.

At test time the callers are copied from the reference checkout at
/root/reference into a temporary directory and edited there; nothing of the
reference is kept in this repository.  Skips when the checkout, g++ or (for
the MPI build) the MPICH headers are absent.

Recipes (each edit names the reference line it changes):
  src_mpi, FP64 ABI (CUDA_REAL = double):
    def.h:69            remove `#define CUDA_FLOAT`
    def.h:80            remove `#define NSIGHT` (NVTX ranges need nvToolsExt.h)
    particle.h:234      guard the CUDA_REAL overload of
                        predictParticleSecondOrder with #ifdef CUDA_FLOAT
    Particle/UpdateParticle.cpp:13-35  guard its definition the same way
    cuda/CalculateRegularAcceleration.cpp:8        include abyss_regf.h
    cuda/CalculateRegularAcceleration.cpp:164-169  guard the float -> double
                        copy loop with #ifdef CUDA_FLOAT
    main.cpp:17         remove `#include <cuda_runtime.h>`
    main.cpp:18         include abyss_regf.h
    main.cpp:55         remove `cudaDeviceSynchronize();` (OpenDevice is
                        synchronous)
  src_mpi, FP32 ABI (CUDA_REAL = float, unchanged def.h:69; link
  libabyss_regf_f32.so, compile with -DABYSS_LEGACY_FLOAT):
    def.h:80, main.cpp:17/18/55 and cuda/CalculateRegularAcceleration.cpp:8
    as above; nothing else.
  src (serial, CUDA_REAL already double, src/defs.h:97-102):
    cuda/CalculateRegularAcceleration.cpp:7, cuda/CalculateAccelerationForAll.cpp:6,
    main.cpp:5          include abyss_regf.h
"""
import os
import re
import shutil
import subprocess

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
MPI_INC = "/opt/conda/include"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="reference checkout or g++ absent")


def _read(p):
    with open(p) as f:
        return f.read().split("\n")


def _write(p, lines):
    with open(p, "w") as f:
        f.write("\n".join(lines))


def _find(lines, pattern, start=0):
    rx = re.compile(pattern)
    for k in range(start, len(lines)):
        if rx.search(lines[k]):
            return k
    raise AssertionError(f"pattern not found: {pattern}")


def remove_line(path, pattern):
    lines = _read(path)
    k = _find(lines, pattern)
    lines[k] = ""
    _write(path, lines)


def replace_line(path, pattern, new):
    lines = _read(path)
    k = _find(lines, pattern)
    lines[k] = new
    _write(path, lines)


def _guard_at(lines, k0, macro):
    depth, k1 = 0, None
    for k in range(k0, len(lines)):
        depth += lines[k].count("{") - lines[k].count("}")
        if depth == 0 and "{" in "".join(lines[k0:k + 1]):
            k1 = k
            break
    assert k1 is not None, f"unbalanced block at line {k0 + 1}"
    return lines[:k0] + [f"#ifdef {macro}"] + lines[k0:k1 + 1] + ["#endif"] + lines[k1 + 1:]


def guard_block(path, start_pattern, macro="CUDA_FLOAT"):
    """Wraps the braced block that opens on the first line matching
    start_pattern (through its closing brace) in #ifdef macro / #endif."""
    lines = _read(path)
    _write(path, _guard_at(lines, _find(lines, start_pattern), macro))


def guard_line(path, pattern, macro="CUDA_FLOAT"):
    lines = _read(path)
    k = _find(lines, pattern)
    lines = lines[:k] + [f"#ifdef {macro}", lines[k], "#endif"] + lines[k + 1:]
    _write(path, lines)


def guard_loop_using(path, use_pattern, macro="CUDA_FLOAT"):
    """Wraps the `for (int i=0; i<ListSize; i++)` loop enclosing the first
    line matching use_pattern."""
    lines = _read(path)
    ku = _find(lines, use_pattern)
    rx = re.compile(r"^\s*for \(int i=0; i<ListSize; i\+\+\) \{")
    k0 = max(k for k in range(ku) if rx.search(lines[k]))
    _write(path, _guard_at(lines, k0, macro))


def syntax(cwd, src, includes, defines=()):
    cmd = ["g++", "-std=c++11", "-fsyntax-only", *[f"-D{d}" for d in defines],
           *[f"-I{i}" for i in includes], src]
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    return p.returncode, p.stderr


MPI_DEFS = ("CUDA", "MPICH_SKIP_MPICXX", "AR_TTL", "AR_SLOWDOWN_TREE", "AR_SLOWDOWN_TIMESCALE",
            "FEWBODY")


def _mpi_tree(tmp_path):
    if not os.path.exists(os.path.join(MPI_INC, "mpi.h")):
        pytest.skip("MPICH headers absent")
    d = tmp_path / "src_mpi"
    shutil.copytree(os.path.join(REF, "src_mpi"), d)
    return str(d)


def _mpi_includes(d):
    return [d, os.path.join(d, "Particle"), os.path.join(d, "cuda"),
            os.path.join(REF, "SDAR", "src"), MPI_INC, INCLUDE]


def _common_mpi_edits(d):
    remove_line(os.path.join(d, "def.h"), r"^#define NSIGHT\b")
    replace_line(os.path.join(d, "cuda", "CalculateRegularAcceleration.cpp"),
                 r'^#include "cuda_functions.h"', '#include "abyss_regf.h"')
    main = os.path.join(d, "main.cpp")
    remove_line(main, r"^#include <cuda_runtime.h>")
    replace_line(main, r'^#include "cuda/cuda_functions.h"', '#include "abyss_regf.h"')
    remove_line(main, r"^\s*cudaDeviceSynchronize\(\);\s*$")


def test_src_mpi_fp64_recipe_compiles(tmp_path):
    d = _mpi_tree(tmp_path)
    # the unedited caller does not build against the FP64 header
    rc, err = syntax(d, "cuda/CalculateRegularAcceleration.cpp", _mpi_includes(d), MPI_DEFS)
    assert rc != 0
    remove_line(os.path.join(d, "def.h"), r"^#define CUDA_FLOAT\s*$")
    guard_line(os.path.join(d, "particle.h"),
               r"void predictParticleSecondOrder\(double dt, CUDA_REAL pos\[\], "
               r"CUDA_REAL vel\[\]\);")
    guard_block(os.path.join(d, "Particle", "UpdateParticle.cpp"),
                r"^void Particle::predictParticleSecondOrder\(double dt, CUDA_REAL pos\[\], "
                r"CUDA_REAL vel\[\]\) \{")
    guard_loop_using(os.path.join(d, "cuda", "CalculateRegularAcceleration.cpp"),
                     r"AccRegReceive_f\[i\]\[dim\];")
    _common_mpi_edits(d)
    for src in ("cuda/CalculateRegularAcceleration.cpp", "main.cpp",
                "Particle/UpdateParticle.cpp"):
        rc, err = syntax(d, src, _mpi_includes(d), MPI_DEFS)
        assert rc == 0, f"{src}:\n{err[-3000:]}"


def test_src_mpi_fp32_recipe_compiles(tmp_path):
    d = _mpi_tree(tmp_path)
    _common_mpi_edits(d)
    for src in ("cuda/CalculateRegularAcceleration.cpp", "main.cpp"):
        rc, err = syntax(d, src, _mpi_includes(d), MPI_DEFS + ("ABYSS_LEGACY_FLOAT",))
        assert rc == 0, f"{src}:\n{err[-3000:]}"


def test_src_serial_recipe_compiles(tmp_path):
    d = str(tmp_path / "src")
    shutil.copytree(os.path.join(REF, "src"), d)
    for f, pat in (("cuda/CalculateRegularAcceleration.cpp", r'^#include "cuda_functions.h"'),
                   ("cuda/CalculateAccelerationForAll.cpp", r'^#include "cuda_functions.h"'),
                   ("main.cpp", r'^#include "cuda/cuda_functions.h"')):
        replace_line(os.path.join(d, f), pat, '#include "abyss_regf.h"')
    for src in ("cuda/CalculateRegularAcceleration.cpp", "cuda/CalculateAccelerationForAll.cpp",
                "main.cpp"):
        rc, err = syntax(d, src, [d, os.path.join(d, "cuda"), INCLUDE], ("CUDA",))
        assert rc == 0, f"{src}:\n{err[-3000:]}"


@pytest.mark.parametrize("tree,defs_header,extra", [
    ("src", "defs.h", ()),
    ("src_mpi", "def.h", ("ABYSS_LEGACY_FLOAT",)),
])
def test_legacy_declarations_match_the_reference(tmp_path, tree, defs_header, extra):
    """The reference's own cuda_functions.h (with its own CUDA_REAL: double in
    src/defs.h, float in the unedited src_mpi/def.h) and abyss_regf.h declare
    the six entry points identically: a mismatch is a conflicting
    declaration of a C function."""
    d = str(tmp_path / tree)
    os.makedirs(os.path.join(d, "cuda"))
    shutil.copy(os.path.join(REF, tree, "cuda", "cuda_functions.h"), os.path.join(d, "cuda"))
    shutil.copy(os.path.join(REF, tree, defs_header), d)
    tu = os.path.join(d, "tu.cpp")
    with open(tu, "w") as f:
        f.write('#include "cuda/cuda_functions.h"\n#include "abyss_regf.h"\n'
                "static_assert(sizeof(abyss_legacy_real) == sizeof(CUDA_REAL), \"real\");\n")
    rc, err = syntax(d, "tu.cpp", [d, INCLUDE], extra)
    assert rc == 0, err[-3000:]
    # and the other scalar type conflicts (the check can fail)
    other = () if extra else ("ABYSS_LEGACY_FLOAT",)
    rc, err = syntax(d, "tu.cpp", [d, INCLUDE], other)
    assert rc != 0 and "conflicting declaration" in err
