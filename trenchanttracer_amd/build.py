"""Builds the native library libtt.so (host C++ with g++, HIP kernels with hipcc
for gfx950) in-tree, next to this file."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtt.so")

HOST_SRC = ["host/sbvh.cpp", "host/flatten.cpp", "host/gen.cpp", "host/api_host.cpp", "host/loaders.cpp"]
DEV_SRC = ["device/trace.hip", "device/api_dev.hip", "device/env.hip", "device/render.hip", "device/guide.hip",
           "device/reorder.hip", "device/display.hip"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# Host code: one IEEE binary32 operation per source operation (no contraction,
# no -march), the arithmetic the reference's host builder performs.
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            "-I", os.path.join(ROOT, "include")]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIPFLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + CXXFLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HOST_SRC + DEV_SRC if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
