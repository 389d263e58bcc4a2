"""Build libf2m_hip.so (gfx950) and the f2m_solver CLI in-tree with hipcc.

    python -m f2m_amd.build            # library + CLI
Objects go to f2m_amd/_build/ (git-ignored); the library is
f2m_amd/libf2m_hip.so, the CLI f2m_amd/f2m_solver.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import re
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


def _asm(src):
    return os.path.join(BUILD, os.path.basename(src) + ".gfx950.s")


_LABEL = re.compile(r"^(\.LBB\S+:|; %bb\.\d+:|_Z\S+:)")
_SPILL = re.compile(r"(Folded Spill|Folded Reload|Spill$|Reload$)")


def lost_spills(asm_text: str):
    """Spill stores/reloads that execute before their block's exec restore.

    A join block of divergent code starts with the narrowed exec mask of the
    branch it ends and restores the full mask with `s_or_b64 exec, exec, ...`.
    A spill placed ahead of that restore runs for the narrowed (often empty)
    set of lanes, so the other lanes' values are lost; the compiler has been
    seen to place spills there.  Returns (kernel, line number, instruction)
    for every such spill.

    Spills inside a whole-wave region are safe: the compiler saves the VGPR
    that holds spilled SGPRs (v_writelane) with every lane on, bracketed by
    `s_or_saveexec_b64 s[..], -1` and `s_mov_b64 exec, s[..]`."""
    out = []
    kernel = None
    block = []
    def flush():
        restore = next((i for i, (_, t) in enumerate(block) if t.startswith("s_or_b64 exec, exec")), None)
        if restore is not None:
            wwm = False
            for ln, t in block[:restore]:
                if t.startswith("s_or_saveexec_b64") and t.replace(" ", "").endswith(",-1"):
                    wwm = True
                elif t.startswith("s_mov_b64 exec,"):
                    wwm = False
                elif not wwm and _SPILL.search(t) and t.startswith(("scratch_", "buffer_", "v_accvgpr")):
                    out.append((kernel, ln, t))
    for ln, raw in enumerate(asm_text.split("\n"), 1):
        line = raw.strip()
        m = _LABEL.match(line)
        if m:
            flush()
            block = []
            if line.startswith("_Z"):
                kernel = line.split(":")[0]
            continue
        if line and not line.startswith((";", ".")):
            block.append((ln, line))
    flush()
    return out


def check_isa(srcs):
    """Fail the build when a device spill can run with lanes masked off."""
    bad = []
    for s in srcs:
        with open(_asm(s)) as f:
            bad += [(os.path.basename(s),) + b for b in lost_spills(f.read())]
    if bad:
        msg = "\n".join(f"  {f}: {k} line {ln}: {t}" for f, k, ln, t in bad[:20])
        raise RuntimeError("device code places spills before an exec restore (values of masked lanes "
                           "would be lost):\n" + msg)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cpp_srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = []
    for s in hip_srcs:
        jobs.append([HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", s, "-o", _obj(s)])
        jobs.append([HIPCC, *COMMON, *DEVICE, "-x", "hip", "--cuda-device-only", "-S", s, "-o", _asm(s)])
    for s in cpp_srcs:
        jobs.append(["g++", *COMMON, "-c", s, "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    check_isa(hip_srcs)
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
