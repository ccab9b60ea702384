"""Builds libstencil_hip.so (HIP, gfx950) and the hstencil3d_hip front end in-tree.

    python -m stencil3d_amd.build          # or stencil3d_amd.build.build()

Every HIP source is compiled with hipcc for gfx950 only, with
-ffp-contract=off (the sweep spells out every fused multiply-add, so the
reference's rounding order is reproduced exactly) and without packed FP32
instructions.  Outputs: stencil3d_amd/libstencil_hip.so and
stencil3d_amd/hstencil3d_hip; objects go to stencil3d_amd/_build/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libstencil_hip.so")
FRONTEND = os.path.join(PKG, "hstencil3d_hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

HIP_SOURCES = ["sth_api.hip", "fields.hip", "sweep_7.hip", "sweep_13.hip", "sweep_25.hip",
               "sweep16.hip", "sweep2.hip", "batch.hip", "noalloc.hip"]


def _run(cmd, quiet):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    if not quiet:
        # hipcc's host pass reports the device-only target feature; drop that line
        msg = "\n".join(line for line in res.stderr.splitlines()
                        if "packed-fp32-ops" not in line and line.strip())
        if msg:
            sys.stderr.write(msg + "\n")


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build(force: bool = True, quiet: bool = True, jobs: int = 8) -> str:
    """Compile every HIP source for gfx950 and link the library + front end.
    force=False skips up-to-date objects.  Returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(INCLUDE, "stencil_hip.h"))
    objs, jobs_list = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs_list.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, quiet), jobs_list))
    if force or jobs_list or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], quiet)
        os.replace(tmp, LIB)
    fe_src = os.path.join(CSRC, "frontend", "hstencil3d_hip.cpp")
    if force or not _newer(FRONTEND, [fe_src, LIB, os.path.join(INCLUDE, "stencil_hip.h")]):
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", fe_src, "-o", FRONTEND,
              f"-L{PKG}", "-lstencil_hip", "-Wl,-rpath,$ORIGIN", "-lm"], quiet)
    return LIB


REFDIR = os.environ.get("STH_REFDIR", "/root/reference")


def build_refbind(quiet: bool = True) -> bool:
    """Where the reference's sources are present: its own driver (main.cpp)
    with csrc/refbind/iterators_hip.cpp in place of iterators.cpp, linked with
    the library, into stencil3d_amd/_refbind/.  Returns whether it was built."""
    if not os.path.isfile(os.path.join(REFDIR, "main.cpp")):
        return False
    subprocess.run(["make", "-C", os.path.join(CSRC, "refbind"), "-j6", f"REFDIR={REFDIR}"],
                   check=True, stdout=subprocess.DEVNULL if quiet else None)
    return True


if __name__ == "__main__":
    build(force="--force" in sys.argv or True, quiet=False)
    print(LIB)
