"""Builds libmxr.so (gfx950) in-tree with hipcc.

    python -m matx_amd.build            # compile every csrc/*.hip, link libmxr.so

Objects go to matx_amd/_build/, the library to matx_amd/libmxr.so.  All HIP
code is compiled with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmxr.so")
ARCH = os.environ.get("MXR_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    "-O3", "-fPIC", "-std=c++20", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj):
        deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + [os.path.join(ROOT, "include", "mxr.h")]
        if os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
            return obj
    cmd = [HIPCC, "-c", src, "-o", obj] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj


CPP_PROGRAMS = {
    # C++ front-end programs linked against libmxr.so.  They are compiled with
    # hipcc because MatX-style user operators (BaseOp) carry device code that
    # is instantiated in the user's translation unit; the built-in operators
    # reach the library through the C ABI either way.
    "simple_radar_pipeline": os.path.join(HERE, "examples", "simple_radar_pipeline.cpp"),
    "radar_stage_test": os.path.join(ROOT, "tests", "cpp", "radar_stage_test.cpp"),
}
BIN = os.path.join(BUILD, "bin")


def _build_one_cpp(name: str, src: str) -> None:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    out = os.path.join(BIN, name)
    cmd = [HIPCC, "-x", "hip", "-std=c++20", "-O2", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
           "-I" + os.path.join(ROOT, "include"), src, "-o", out, "-L" + HERE, "-lmxr",
           "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")


def build_cpp(verbose: bool = True) -> None:
    os.makedirs(BIN, exist_ok=True)
    with cf.ThreadPoolExecutor(len(CPP_PROGRAMS)) as ex:
        list(ex.map(lambda kv: _build_one_cpp(*kv), CPP_PROGRAMS.items()))
    if verbose:
        print(f"built {len(CPP_PROGRAMS)} C++ front-end programs in {BIN}")


def build(verbose: bool = True, jobs: int | None = None, incremental: bool = False) -> str:
    """Compiles every csrc/*.hip (all of them unless `incremental`, which
    skips objects newer than their source and the shared headers) and links
    libmxr.so, then the C++ front-end programs."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), srcs))
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-L" + rocm_lib, "-lhiprtc", "-Wl,-rpath," + rocm_lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_cpp(verbose)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
