"""Builds the gfx950 library and the C++ host programs in-tree.

    python -m cfd_optimisations_amd.build [--incremental] [--jobs N]

Products (git-ignored, shipped with the tree):
    cfd_optimisations_amd/libhcfd.so       C ABI + HIP kernels (include/hcfd.h)
    cfd_optimisations_amd/hcfd_run         C++ drop-in driver (namespace MI355X classes)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhcfd.so")
RUN = os.path.join(PKG, "hcfd_run")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
# -ffp-contract=off: the FP64 path keeps the reference's rounding step by step.
# Packed FP32 instructions are disabled for all device code: some of their
# forms have given silently wrong results on this hardware.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I" + INCLUDE, "-I" + CSRC]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE, "-I" + HOST]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libhcfd.so")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def _newest(paths):
    return max((os.path.getmtime(p) for p in paths), default=0.0)


def build_library(incremental: bool = False, jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    hdr_time = _newest(headers)
    cc = hipcc()
    jobs_todo = []
    objs = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        if incremental and os.path.exists(o) and os.path.getmtime(o) > max(os.path.getmtime(s), hdr_time):
            continue
        cmd = [cc] + HIP_FLAGS + os.environ.get("HCFD_EXTRA_HIP_FLAGS", "").split() + ["-c", s, "-o", o]
        if s.endswith(".cpp"):
            cmd = [cc] + HIP_FLAGS + ["-x", "c++", "-c", s, "-o", o]
        jobs_todo.append(cmd)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, jobs_todo):
            if verbose and out.strip():
                print(out)
    if not incremental or not os.path.exists(LIB) or os.path.getmtime(LIB) < _newest(objs):
        tmp = LIB + ".tmp"
        _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    return LIB


def build_host(incremental: bool = False) -> str:
    """The C++ drop-in classes (namespace MI355X) and the Test.cpp-style driver."""
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not srcs:
        return ""
    if incremental and os.path.exists(RUN) and os.path.getmtime(RUN) > max(
            _newest(srcs + glob.glob(os.path.join(HOST, "*.hpp")) + glob.glob(os.path.join(HOST, "MI355X", "*.hpp"))),
            os.path.getmtime(LIB)):
        return RUN
    cxx = shutil.which("g++") or "g++"
    _run([cxx] + HOST_FLAGS + ["-o", RUN] + srcs + ["-L" + PKG, "-lhcfd", "-Wl,-rpath,$ORIGIN"])
    return RUN


def build_checkers() -> None:
    """The oracle (always) and, when the reference sources are present, its own build."""
    od = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", od, "oracle"])
    if os.path.isdir(os.environ.get("HCFD_REFDIR", "/root/reference")):
        _run(["make", "-s", "-C", od, "ref", "REFDIR=" + os.environ.get("HCFD_REFDIR", "/root/reference")])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-checkers", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    build_library(a.incremental, a.jobs, a.verbose)
    build_host(a.incremental)
    if not a.no_checkers:
        build_checkers()
    print("built", LIB)
    return 0


if __name__ == "__main__":
    sys.exit(main())
