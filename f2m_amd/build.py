"""Build libf2m_hip.so (gfx950) and the f2m_solver CLI in-tree with hipcc.

    python -m f2m_amd.build            # library + CLI
Objects go to f2m_amd/_build/ (git-ignored); the library is
f2m_amd/libf2m_hip.so, the CLI f2m_amd/f2m_solver.
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
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libf2m_hip.so")
CLI = os.path.join(PKG, "f2m_solver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cpp_srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = []
    for s in hip_srcs:
        jobs.append([HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", s, "-o", _obj(s)])
    for s in cpp_srcs:
        jobs.append(["g++", *COMMON, "-c", s, "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    objs = [_obj(s) for s in hip_srcs + cpp_srcs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli_src = os.path.join(PKG, "cli", "f2m_solver.cpp")
    if os.path.exists(cli_src):
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src, "-o", CLI,
              "-L", PKG, "-lf2m_hip", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
