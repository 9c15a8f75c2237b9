"""Builds the native libraries in-tree.

libxdem_hip.so    HIP kernels + C ABI (hipcc, gfx950 only)
libxdem_host.so   host C++: input generators, operator registry and .msp deck runner (g++)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")

HIP_SOURCES = [
    "api.hip",
    "integrate.hip",
    "grid.hip",
    "contact_sphere.hip",
    "contact_sphere_fp32.hip",
    "poly.hip",
    "gnn.hip",
    "analysis.hip",
    "mesh.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# FP64 plain form: no floating-point contraction anywhere, so that results reproduce the CPU
# restatement bit for bit; packed FP32 instructions are disabled for every kernel.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, "libxdem_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out, *objs])
    for o in objs:
        os.remove(o)
    return out


HOST_SOURCES = [
    "host/generators.cpp",
    "host/msp.cpp",
    "host/deck.cpp",
]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]


def build_host():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libxdem_host.so")
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    _run([CXX, *HOST_FLAGS, "-shared", "-o", out, *srcs,
          "-L" + LIBDIR, "-lxdem_hip", "-Wl,-rpath,$ORIGIN"])
    exe = os.path.join(LIBDIR, "xdem-run")
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-o", exe, os.path.join(CSRC, "host", "xdem_run.cpp"),
          "-L" + LIBDIR, "-lxdem_host", "-lxdem_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    return build_hip(), build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
