"""Build libfvh.so (HIP, gfx950) in-tree with hipcc.

    python finite-volume-2d-cpp_amd/build.py [--arch gfx950] [-j N]

All sources compile with FP contraction off (the FP64 path must round like
the reference's x86-64 build) and without packed-FP32 instructions.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfvh.so")
SOURCES = ["fvh_mesh.cpp", "fvh_kernels.hip", "fvh_adjoint.hip", "fvh_solver.hip"]
DROPIN = os.path.join(HERE, "dropin")
DROPIN_SOURCES = ["fvhyper_dropin.cpp", "probe.cpp"]
DROPIN_LIB = os.path.join(HERE, "libfvhyper.so")
CHECK_LIB = os.path.join(HERE, "libfvh_check.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def flags(arch: str):
    return [
        "--offload-arch=%s" % arch,
        "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        "-I%s" % CSRC, "-I%s" % os.path.join(os.path.dirname(HERE), "include"),
    ]


def compile_one(src: str, arch: str, srcdir: str = CSRC, extra=()) -> str:
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC] + flags(arch) + list(extra) + ["-c", os.path.join(srcdir, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + flags(arch) + list(extra) + ["-x", "hip", "-c", os.path.join(srcdir, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(arch: str = "gfx950", jobs: int = 4) -> str:
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, arch), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=%s" % arch, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # the fvhyper drop-in layer (C++ headers with the reference's API)
    inc = ["-I%s" % os.path.join(DROPIN, "include")]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        dobjs = list(ex.map(lambda s: compile_one(s, arch, os.path.join(DROPIN, "src"), inc), DROPIN_SOURCES))
    tmp = DROPIN_LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=%s" % arch, "-shared", "-fPIC", "-o", tmp] + dobjs + \
          ["-L%s" % HERE, "-lfvh", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, DROPIN_LIB)
    # host-only checks of the physics and its reverse (CPU tests)
    cobj = compile_one("fvh_check.cpp", arch)
    tmp = CHECK_LIB + ".tmp"
    mobj = os.path.join(OUT, "fvh_mesh.cpp.o")   # host mesh code, for the geometry reverse
    r = subprocess.run([HIPCC, "--offload-arch=%s" % arch, "-shared", "-fPIC", "-o", tmp, cobj, mobj],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, CHECK_LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args()
    print(build(a.arch, a.j))
    sys.exit(0)
