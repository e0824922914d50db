"""Build recipe of the HIP library (gfx950) and of the test oracle.

    python -m strassen_amd.build          # library + oracle (+ reference shim if present)

The library is compiled in-tree with hipcc into strassen_amd/lib/libstrassen_amd.so
(one object per .hip source, then one shared link).  All HIP code is built with
``-Xclang -target-feature -Xclang -packed-fp32-ops``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libstrassen_amd.so")
ARCH = os.environ.get("SA_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + [os.path.join(ROOT, "include", "strassen_amd.h")]
    hdr_mtime = max(os.path.getmtime(h) for h in headers)

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        if os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), hdr_mtime):
            return obj
        out = _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lrocblas"])
    os.replace(tmp, LIB)
    return LIB


CPP = os.path.join(PKG, "cpp")
BIN = os.path.join(PKG, "bin")
DRIVERS = ["occ_ri_k_step", "3c_2c_transform", "strassen_bench"]


def build_cpp() -> list:
    """Drop-in C++ drivers (reference argv) linked against the HIP library."""
    os.makedirs(BIN, exist_ok=True)
    out = []

    def one(src):
        exe = os.path.join(BIN, os.path.splitext(os.path.basename(src))[0])
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", CPP,
              src, "-L", LIBDIR, "-lstrassen_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe])
        return exe

    # the reference-argv drivers, and the C++ programs the GPU tests drive
    srcs = [os.path.join(CPP, "drivers", name + ".cpp") for name in DRIVERS]
    srcs += sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        out = list(ex.map(one, srcs))
    return out


def build_oracle(with_ref: bool = True) -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "all"])
    if with_ref and os.path.isdir(os.environ.get("SA_REFERENCE_DIR", "/root/reference")):
        try:
            _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
        except RuntimeError as e:  # the reference build is optional (needs MKL)
            print(f"[build] reference shim not built: {e}", file=sys.stderr)


if __name__ == "__main__":
    build_oracle()
    print(build_library(verbose=True))
    print(build_cpp())
