"""Compiles libcafe_gpu.so (the HIP engine) in-tree for gfx950.

    python -m cafe5_amd.build          # builds cafe5_amd/libcafe_gpu.so

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled for every HIP file (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcafe_gpu.so")
SOURCES = ["api.hip", "build_kernels.hip", "prune_kernels.hip", "prune_mfma.hip", "prune_lane.hip", "prune_f16.hip",
           "reconstruct.hip"]
ARCH = os.environ.get("CAFE_GPU_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


# The plan stress driver (tests/cpp/plan_stress.cpp): many plans of every path
# alive at once through the C ABI (a GPU test runs it), built twice: against
# libcafe_gpu.so, and against a copy whose host code is compiled with
# AddressSanitizer (device code is not instrumented).  The instrumented copy is
# linked without a sanitizer runtime; the driver brings GCC's (g++
# -fsanitize=address), which instruments the host heap and stack only -- clang's
# runtime also intercepts the HSA runtime's device allocations, which this
# image's runtime does not support without XNACK.
STRESS_DRIVER = os.path.join(ROOT, "tests", "cpp", "plan_stress")
ASAN_DIR = os.path.join(ROOT, "tests", "cpp", "_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libcafe_gpu.so")
ASAN_DRIVER = os.path.join(ASAN_DIR, "plan_stress")
ASAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g1"]


def _compile_all(variants, verbose):
    """variants: [(extra flags, object directory)]; compiles every source of
    every variant in parallel; returns the object lists."""
    hip = hipcc()
    flags = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
    procs, objs = [], []
    for extra, obj_dir in variants:
        os.makedirs(obj_dir, exist_ok=True)
        objs.append([])
        for src in SOURCES:
            obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
            objs[-1].append(obj)
            cmd = [hip, *flags, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    return objs


def _run(cmd, what):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"failed to build {what}")


def _link(objs, lib, extra, verbose):
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *extra, *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed to link {lib}")
    os.replace(tmp, lib)


def build_native(verbose: bool = False, asan: bool = False) -> str:
    """Compile every HIP source (in parallel) and link cafe5_amd/libcafe_gpu.so;
    with asan, also the host-instrumented copy and its stress driver under
    tests/cpp/_asan (compiled in the same parallel batch).  Returns the
    library's path."""
    variants = [([], os.path.join(HERE, "_build"))]
    if asan:
        variants.append((ASAN_FLAGS, os.path.join(HERE, "_build_asan")))
    objs = _compile_all(variants, verbose)
    _link(objs[0], LIB, [], verbose)
    if asan:
        os.makedirs(ASAN_DIR, exist_ok=True)
        _link(objs[1], ASAN_LIB, [], verbose)
        tmp = ASAN_DRIVER + ".tmp"
        cmd = ["g++", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-std=c++17",
               "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "plan_stress.cpp"),
               "-o", tmp, "-L", ASAN_DIR, "-lcafe_gpu", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd, "tests/cpp/_asan/plan_stress")
        os.replace(tmp, ASAN_DRIVER)
    return LIB


HOST_LIB = os.path.join(HERE, "libcafe5_host.so")
ADAPTER_TEST = os.path.join(ROOT, "tests", "cpp", "adapter_test")


def build_host(verbose: bool = False) -> str:
    """The C++ host interface (cafe5_amd/host) over libcafe_gpu.so, and the
    C++ test driver that exercises it (tests/cpp/adapter_test)."""
    src = os.path.join(HERE, "host", "cafe5_gpu.cpp")
    common = ["-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host")]
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", *common, "-fPIC", "-shared", src, "-o", tmp, "-L", HERE, "-lcafe_gpu",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, "libcafe5_host.so")
    os.replace(tmp, HOST_LIB)
    test_src = os.path.join(ROOT, "tests", "cpp", "adapter_test.cpp")
    tmp = ADAPTER_TEST + ".tmp"
    cmd = ["g++", *common, test_src, "-o", tmp, "-L", HERE, "-lcafe5_host", "-lcafe_gpu",
           "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(ADAPTER_TEST)).join(["$ORIGIN/", ""])]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, "tests/cpp/adapter_test")
    os.replace(tmp, ADAPTER_TEST)
    tmp = STRESS_DRIVER + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "plan_stress.cpp"), "-o", tmp, "-L", HERE, "-lcafe_gpu",
           "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(STRESS_DRIVER)).join(["$ORIGIN/", ""])]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, "tests/cpp/plan_stress")
    os.replace(tmp, STRESS_DRIVER)
    return HOST_LIB


if __name__ == "__main__":
    print(build_native(verbose=True, asan="--asan" in sys.argv))
    print(build_host(verbose=True))
