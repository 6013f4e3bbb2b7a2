"""Build recipe of liblepic_hip.so (gfx950 only) with hipcc.

The FP64 hot path relies on strict IEEE arithmetic, so everything is built
with -ffp-contract=off (mixed-precision kernels request fused multiply-adds
explicitly with fmaf).  Packed FP32 instructions are disabled for all HIP
code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblepic_hip.so")
SOURCES = ["lepic_ctx.hip", "push.hip", "field.hip", "poisson.hip", "sort.hip", "mcc.hip", "loader.hip",
           "unet.hip", "checkpoint.hip", "moments.hip", "heating.hip", "sources.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
         "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]


def build(verbose: bool = False, jobs: int = 8) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for p in procs if p[1].poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if p.returncode != 0 or (verbose and lines):
            print("\n".join(lines), file=sys.stderr)
        if p.returncode != 0:
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran_driver(objdir)
    return LIB


def build_fortran_driver(objdir: str) -> str | None:
    """The Fortran time loop over the C ABI (iso_c_binding), when amdflang is
    present; links liblepic_hip.so from the package directory."""
    flang = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
    if not os.path.exists(flang):
        return None
    fdir = os.path.join(HERE, "fortran")
    exe = os.path.join(fdir, "lepic_driver")
    cmd = [flang, "-O2", "-module-dir", objdir, os.path.join(fdir, "lepic_capi.f90"),
           os.path.join(fdir, "lepic_driver.f90"), "-L" + HERE, "-llepic_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", exe]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
