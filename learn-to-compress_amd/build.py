"""Builds libleco_fix.so (gfx950 HIP kernels + C ABI) in-tree.

    python learn-to-compress_amd/build.py

Every HIP source is compiled with -ffp-contract=off: the kernels place their
fused multiply-adds explicitly (see csrc/leco_device.h), and contraction by the
compiler would change theta and prediction bits.  Packed FP32 instructions are
disabled (-packed-fp32-ops), none of the kernels wants them.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libleco_fix.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["leco_kernels.hip", "leco_filter.hip", "leco_api.hip", "leco_batch.hip", "leco_serve.hip", "leco_io.hip",
           "leco_base.hip", "leco_opmax.hip"]
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}")
    if out:
        print(out, file=sys.stderr)


def build() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "leco_int_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "leco_int_driver")


BASE_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "base_driver.cpp")
BASE_DRIVER = os.path.join(ROOT, "tests", "cpp", "base_driver")
OPMAX_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "opmax_driver.cpp")
OPMAX_DRIVER = os.path.join(ROOT, "tests", "cpp", "opmax_driver")
CHAIN_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "chain_check.cpp")
CHAIN_CHECK = os.path.join(ROOT, "tests", "cpp", "chain_check")
PERIOD_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "period_check.cpp")
PERIOD_CHECK = os.path.join(ROOT, "tests", "cpp", "period_check")


def build_driver() -> str:
    """The C++ drivers over include/leco_int.hpp, leco_baselines.hpp and
    leco_opmax.hpp (host code: g++ links the library), and the host-side checks
    of the op_max chain evaluation and of the close filter's period index
    (leco_chain.h and leco_period.h are HIP host+device code, so hipcc compiles
    them; only their host code runs)."""
    for src, exe in ((DRIVER_SRC, DRIVER), (BASE_DRIVER_SRC, BASE_DRIVER), (OPMAX_DRIVER_SRC, OPMAX_DRIVER)):
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), src,
              "-o", exe, "-L", LIBDIR, "-lleco_fix", "-Wl,-rpath,$ORIGIN/../../learn-to-compress_amd/lib"])
    for src, exe in ((CHAIN_CHECK_SRC, CHAIN_CHECK), (PERIOD_CHECK_SRC, PERIOD_CHECK)):
        _run([HIPCC, "-x", "hip", f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
              src, "-o", exe])
    return DRIVER


if __name__ == "__main__":
    print(build())
