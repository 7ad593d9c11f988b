"""Build the HIP libraries in-tree (gfx950 only; host-only sources with g++):
  abyss_amd/libabyss_regf.so      legacy entry points with double (CUDA_REAL = double)
  abyss_amd/libabyss_regf_f32.so  the same library with float legacy entry points
                                  (regf_legacy.hip built with -DABYSS_LEGACY_FLOAT)

Every HIP source is compiled with packed-FP32 instruction selection turned
off (-Xclang -target-feature -Xclang -packed-fp32-ops).  Both libraries bind
their own symbols first (-Bsymbolic), so the two can be loaded side by side.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libabyss_regf.so")
LIB_F32 = os.path.join(HERE, "libabyss_regf_f32.so")
SOURCES = ["regf_kernels.hip", "regf_runtime.hip", "regf_ensemble.hip", "regf_autotune.hip",
           "regf_state.hip", "regf_irregular.hip", "regf_derivs.hip"]
# (object name, extra flags) of the legacy entry points, one per library
LEGACY = [("regf_legacy.o", []), ("regf_legacy_f32.o", ["-DABYSS_LEGACY_FLOAT"])]
# host-only C++ (no device code), compiled with the host compiler
HOST_SOURCES = ["regf_pack.cpp"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall"]
ARCH = os.environ.get("ABYSS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build(verbose: bool = False, extra=()) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = [(os.path.join(CSRC, s), os.path.join(bdir, s.replace(".hip", ".o")), [])
            for s in SOURCES]
    jobs += [(os.path.join(CSRC, "regf_legacy.hip"), os.path.join(bdir, o), f)
             for o, f in LEGACY]
    procs = []
    for src, o, flags in jobs:
        cmd = [_hipcc(), *HIP_FLAGS, *extra, *flags, "-c", src, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
    host_objs = []
    for src in HOST_SOURCES:
        o = os.path.join(bdir, src.replace(".cpp", ".o"))
        cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        host_objs.append(o)
    for p, cmd in procs:
        if p.wait() != 0:
            raise RuntimeError("build failed: " + " ".join(cmd))
    objs = [o for _, o, _ in jobs[:len(SOURCES)]] + host_objs
    for lib, (legacy_obj, _) in ((LIB, LEGACY[0]), (LIB_F32, LEGACY[1])):
        tmp = lib + ".tmp"
        cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-Bsymbolic",
               "-pthread", "-o", tmp, *objs, os.path.join(bdir, legacy_obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
