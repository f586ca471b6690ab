"""Builds the in-tree libraries for gfx950.

  lib/libfeat3amd.so       HIP kernels + the C-ABI of include/feat3amd.h (hipcc)
  lib/libfeat3amd_host.so  C++ host layer (include/feat3amd/*.hpp) + include/feat3amd_host.h (g++)

Usage: python feat3.doxytest_amd/build.py [--jobs N]
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# every HIP translation unit (one object each, linked into one .so)
HIP_SOURCES = ["feat3amd_kernels.hip", "feat3amd_stokes.hip", "feat3amd_batch.hip", "feat3amd_mlp.hip",
               "feat3amd_assembly.hip", "feat3amd_smoothers.hip", "feat3amd_transpose.hip"]
HOST_SOURCES = ["host_capi.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off: plain float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no implicit contraction: every fused multiply-add is written out
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + INC,
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-I" + INC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {os.path.basename(cmd[-1])}")
    return r.stdout


def build(jobs=4, verbose=False):
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)

    def hip_obj(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(hip_obj, HIP_SOURCES))
    kern = os.path.join(LIB, "libfeat3amd.so")
    tmp = kern + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, kern)

    host = os.path.join(LIB, "libfeat3amd_host.so")
    tmp = host + ".tmp"
    _run([CXX] + HOST_FLAGS + ["-shared", "-o", tmp] +
         [os.path.join(CSRC, s) for s in HOST_SOURCES] +
         ["-L" + LIB, "-lfeat3amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, host)
    shutil.rmtree(objdir, ignore_errors=True)
    if verbose:
        print("built", kern, host)
    return kern, host


if __name__ == "__main__":
    j = 4
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs=j, verbose=True)
