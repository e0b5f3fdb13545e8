"""Builds libdndsr_amd.so (HIP kernels + C-ABI) for gfx950 with hipcc, in-tree."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdndsr_amd.so")
SOURCES = ["setup.hip", "vr.hip", "euler.hip", "adjoint.hip", "linsolve.hip", "model.hip", "roejac.hip", "probe.hip", "capi.hip"]
ARCH = os.environ.get("DNDSR_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 instructions are kept out of every kernel
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every HIP source and link libdndsr_amd.so.  Always recompiles."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


MIRROR_TEST_SRC = os.path.join(HERE, "..", "tests", "cpp", "host_mirror_main.cpp")
MIRROR_TEST_BIN = os.path.join(HERE, "_bin", "host_mirror_test")


def build_host_mirror_test() -> str:
    """Compile the C++17 host-mirror test driver (include/dndsr_amd.hpp) against the library."""
    os.makedirs(os.path.dirname(MIRROR_TEST_BIN), exist_ok=True)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(rocm, "include"), MIRROR_TEST_SRC,
           "-o", MIRROR_TEST_BIN + ".tmp", "-L", HERE, "-ldndsr_amd", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("host mirror test driver failed to build:\n" + r.stdout)
    os.replace(MIRROR_TEST_BIN + ".tmp", MIRROR_TEST_BIN)
    return MIRROR_TEST_BIN


def _wait(procs):
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{out.decode(errors='replace')}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host_mirror_test())
