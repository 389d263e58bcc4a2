"""Known-answer runner over a directory of TSPLIB files (counterpart of the
reference's run_dir.py:1-87).

    python -m f2m_amd.cli.run_dir <dir> <K> [--out DIR] [--known FILE]
        [--max-size N] [--solver PATH] [-- <f2m_solver flags>]

Runs the f2m_solver CLI (counterpart of F2M.solver.exe, RUN.cpp:12-31) on
every `*.tsp` file of <dir> with top-K = <K>; a `*.tsp.gz` file is unpacked
into a temporary directory first (the CLI writes its outputs next to its
input, as the reference does).  From each `<file>.top<K>.F2M_report` it
takes the `Time:`, `Optimal Value:` and `ERROR:` lines, as the reference's
parse_file does (run_dir.py:11-23), and writes, in --out (default: the
current directory):
  report_dir.txt        "<name> Time: <t> Optimal Value: <v>" per solved file
  report_dir_error.txt  "<name> ERROR: Optimal Solution not Found" per failure
(the reference's two files, run_dir.py:58-85).  With --known FILE (lines in
the report_dir.txt format, e.g. a published report_dir.txt) it also writes
known_answers.txt: per instance the published value, ours and their relative
difference, and returns exit code 1 if an instance of FILE is missing.
--max-size N skips files whose name's first number exceeds N (the reference
skips above 100000 by that rule, run_dir.py:43-52; default: no limit).
"""
from __future__ import annotations

import argparse
import gzip
import os
import re
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_SOLVER = os.path.join(PKG, "f2m_solver")


def parse_report(path):
    """(time line, optimal value line, error line) of a .F2M_report; missing
    lines are None (run_dir.py:11-23 keys on the same three markers)."""
    t = v = e = None
    with open(path) as f:
        for line in f:
            line = line.replace("\n", "").replace("\t", "")
            if "Time:" in line:
                t = line
            if "ERROR:" in line:
                e = line
            if "Optimal Value:" in line:
                v = line
    return t, v, e


def read_known(path):
    """{name: value} from lines "<name> Time: <t> Optimal Value: <v>"."""
    out = {}
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*(\S+).*Optimal Value:\s*(\S+)", line)
            if m:
                out[m.group(1)] = float(m.group(2))
    return out


def instance_name(fname):
    for suf in (".tsp.gz", ".tsp"):
        if fname.endswith(suf):
            return fname[: -len(suf)]
    return fname


def size_from_name(fname):
    m = re.search(r"\d+", fname)
    return int(m.group(0)) if m else None


def run(directory, K, out_dir=".", known=None, max_size=None, solver=DEFAULT_SOLVER, flags=(), log=print):
    files = sorted(f for f in os.listdir(directory)
                   if os.path.isfile(os.path.join(directory, f)) and (f.endswith(".tsp") or f.endswith(".tsp.gz")))
    times, values, errors = {}, {}, {}
    tmp = tempfile.mkdtemp(prefix="f2m_run_dir_")
    try:
        for f in files:
            size = size_from_name(f)
            if max_size is not None and size is not None and size > max_size:
                continue
            name = instance_name(f)
            src = os.path.join(directory, f)
            if f.endswith(".gz"):
                path = os.path.join(tmp, name + ".tsp")
                with gzip.open(src, "rb") as fi, open(path, "wb") as fo:
                    shutil.copyfileobj(fi, fo)
            else:
                path = src
            cmd = [solver, path, str(K), *flags]
            log(" ".join(cmd))
            r = subprocess.run(cmd, capture_output=True, text=True)
            rep = f"{path}.top{K}.F2M_report"
            if r.returncode != 0 or not os.path.exists(rep):
                errors[name] = f"ERROR: solver exit code {r.returncode}: {r.stderr.strip()[:200]}"
                continue
            t, v, e = parse_report(rep)
            if e:
                errors[name] = e
            if t and v:
                times[name] = t
                values[name] = v
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "report_dir.txt"), "w") as fo:
        for name in times:
            fo.write(f"{name} {times[name]} {values[name]}\n")
    with open(os.path.join(out_dir, "report_dir_error.txt"), "w") as fo:
        for name, e in errors.items():
            fo.write(f"{name} {e}\n")
    rc = 0
    if known:
        pub = read_known(known)
        with open(os.path.join(out_dir, "known_answers.txt"), "w") as fo:
            fo.write(f"# K={K} flags={' '.join(flags) or '(none)'}\n")
            fo.write("instance published ours rel_diff\n")
            for name, pv in pub.items():
                if name in values:
                    ov = float(values[name].split(":", 1)[1])
                    fo.write(f"{name} {pv!r} {ov!r} {(ov - pv) / abs(pv):.3e}\n")
                else:
                    fo.write(f"{name} {pv!r} missing -\n")
                    rc = 1
    return rc, times, values, errors


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    flags = []
    if "--" in argv:
        i = argv.index("--")
        argv, flags = argv[:i], argv[i + 1:]
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("dir")
    ap.add_argument("K", type=int)
    ap.add_argument("--out", default=".")
    ap.add_argument("--known", default=None)
    ap.add_argument("--max-size", type=int, default=None)
    ap.add_argument("--solver", default=DEFAULT_SOLVER)
    a = ap.parse_args(argv)
    rc, times, values, errors = run(a.dir, a.K, a.out, a.known, a.max_size, a.solver, flags)
    for name in times:
        print(name, " ", times[name], " ", values[name])
    for name, e in errors.items():
        print(name, " ", e)
    return rc


if __name__ == "__main__":
    sys.exit(main())
