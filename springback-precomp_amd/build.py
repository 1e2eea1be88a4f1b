"""Builds libsparlab_gpu.so for gfx950 in-tree (springback-precomp_amd/lib/).

hipcc compiles every HIP and C++ source of csrc/ with
  --offload-arch=gfx950 -O3 -std=c++17 -fPIC
  -Xclang -target-feature -Xclang -packed-fp32-ops
and links them into one shared library that exports the C ABI of
include/sparlab_gpu.h. Sources compile in parallel; every call rebuilds.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsparlab_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as pool:
        objs = list(pool.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


ADAPTER_TEST = os.path.join(ROOT, "tests", "cpp", "_build", "adapter_sequence")


def build_adapter_test(verbose=False):
    """The C++ driver of tests/cpp/adapter_sequence.cpp (the forming driver's
    calls through include/sparlab_gpu_system.hpp), linked to the library."""
    os.makedirs(os.path.dirname(ADAPTER_TEST), exist_ok=True)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           os.path.join(ROOT, "tests", "cpp", "adapter_sequence.cpp"), "-o", ADAPTER_TEST,
           "-L", LIBDIR, "-lsparlab_gpu", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../../springback-precomp_amd/lib", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"the adapter test driver did not build:\n{res.stderr}")
    if verbose:
        print(f"built {ADAPTER_TEST}")
    return ADAPTER_TEST


MATH_PROBE = os.path.join(ROOT, "tests", "cpp", "_build", "libdevice_math_probe.so")


def build_math_probe(verbose=False):
    """The test probe of tests/cpp/device_math_probe.hip (the device math
    helpers beside the IEEE operations), with the library's device flags."""
    os.makedirs(os.path.dirname(MATH_PROBE), exist_ok=True)
    tmp = MATH_PROBE + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", "-o", tmp, os.path.join(ROOT, "tests", "cpp", "device_math_probe.hip")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"the device math probe did not build:\n{res.stderr}")
    os.replace(tmp, MATH_PROBE)
    if verbose:
        print(f"built {MATH_PROBE}")
    return MATH_PROBE


if __name__ == "__main__":
    build(verbose=True)
    build_adapter_test(verbose=True)
    build_math_probe(verbose=True)
    sys.exit(0)
