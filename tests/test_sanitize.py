"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds with g++ -fsanitize=address,undefined -fno-sanitize-recover=all:
  * the TSPLIB reader (f2m_amd/csrc/f2m_tsplib.cpp) and the CPU oracle
    (oracle/f2m_oracle.cpp, oracle/f2m_exact_oracle.cpp) into a driver
    (tests/sanitize/san_driver.cpp) run on TSPLIB files: two golden
    instances and hand-written edge cases (CRLF, duplicates, -0.0, a
    malformed and a missing file);
  * the CLI's host code (f2m_amd/cli/f2m_solver.cpp), linked with the HIP
    library, run on the paths that end before any device call (usage, a
    malformed file, too few points).
Any sanitizer report fails the run."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _files(tmp):
    out = []
    for inst in ("xqf131", "dka1376"):
        p = os.path.join(tmp, inst + ".tsp")
        with gzip.open(os.path.join(ROOT, "tests", "golden", inst + ".tsp.gz"), "rb") as fi, open(p, "wb") as fo:
            fo.write(fi.read())
        out.append(p)
    cases = {
        "crlf.tsp": "NAME : t\r\nNODE_COORD_SECTION\r\n" + "".join(f"{i + 1} {(i * 37) % 101}.25 {(i * 53) % 97}\r\n"
                                                                  for i in range(40)) + "EOF\r\n",
        "dups.tsp": "NODE_COORD_SECTION\n1 0.0 5\n2 -0.0004 5\n3 7 -0.0002\n4 7 0\n" +
                    "".join(f"{i + 5} {i % 9} {i // 9}\n" for i in range(30)) + "5 0 0\nEOF\n",
        "bad.tsp": "NAME : x\nTYPE : TSP\n1 2 3\nEOF\n",
    }
    for name, text in cases.items():
        p = os.path.join(tmp, name)
        with open(p, "w", newline="") as f:
            f.write(text)
        out.append(p)
    out.append(os.path.join(tmp, "missing.tsp"))
    return out


def test_oracle_and_reader_under_sanitizers(tmp_path):
    exe = str(tmp_path / "san_driver")
    cmd = ["g++", "-std=c++17", "-ffp-contract=off", "-pthread", *SAN, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "sanitize", "san_driver.cpp"),
           os.path.join(ROOT, "f2m_amd", "csrc", "f2m_tsplib.cpp"), os.path.join(ROOT, "oracle", "f2m_oracle.cpp"),
           os.path.join(ROOT, "oracle", "f2m_exact_oracle.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    files = _files(str(tmp_path))
    r = subprocess.run([exe, str(tmp_path / "dup_out")] + files, capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.count("objective") == 4  # xqf131, dka1376, crlf, dups
    assert "bad.tsp: status -3" in r.stdout and "missing.tsp: status -4" in r.stdout


def test_cli_host_code_under_sanitizers(tmp_path):
    lib = os.path.join(ROOT, "f2m_amd", "libf2m_hip.so")
    if not os.path.exists(lib):
        pytest.skip("HIP library not built")
    exe = str(tmp_path / "f2m_solver_san")
    cmd = ["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "f2m_amd", "cli", "f2m_solver.cpp"), "-o", exe, "-L", os.path.join(ROOT, "f2m_amd"),
           "-lf2m_hip", "-Wl,-rpath," + os.path.join(ROOT, "f2m_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(ENV, ASAN_OPTIONS="detect_leaks=0")  # the HIP runtime's own allocations are not ours
    bad = tmp_path / "bad.tsp"
    bad.write_text("NAME : x\n1 2 3\nEOF\n")
    few = tmp_path / "few.tsp"
    few.write_text("NODE_COORD_SECTION\n1 0 0\n2 1 0\n3 0 1\nEOF\n")
    for args, want in (([], "usage"), ([str(bad), "10"], "wrong format"), ([str(few), "2"], "at least 4"),
                       ([str(few), "0"], "positive integer")):
        r = subprocess.run([exe] + args, capture_output=True, text=True, env=env, timeout=120)
        assert r.returncode == 0 and want in r.stdout, (args, r.stdout, r.stderr[-2000:])
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
