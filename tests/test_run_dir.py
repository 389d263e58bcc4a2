"""CPU tests of the known-answer runner (counterpart of run_dir.py:1-87): the
directory walk, .tsp.gz unpacking, report parsing, the reference's size
filter and the known-answer table, with a stand-in solver script that writes
reports (no GPU)."""
import gzip
import os
import stat
import sys

from f2m_amd.cli import run_dir

FAKE = """#!{py}
import sys
path, K = sys.argv[1], sys.argv[2]
name = path.rsplit("/", 1)[-1]
with open(path + ".top" + K + ".F2M_report", "w") as f:
    if name.startswith("bad"):
        f.write("ERROR: Optimal Solution not Found\\n")
    else:
        n = sum(1 for ln in open(path) if ln[:1].isdigit())
        f.write("Optimal Found: Yes\\nOptimal Value: %.15g\\nTime: 0.5\\nMAX edge rank: 3\\n" % (n * 1.5))
"""


def _tsp(n):
    return "NAME : t\nNODE_COORD_SECTION\n" + "".join(f"{i + 1} {i} {2 * i}\n" for i in range(n)) + "EOF\n"


def test_run_dir_tables(tmp_path):
    solver = tmp_path / "fake_solver"
    solver.write_text(FAKE.format(py=sys.executable))
    solver.chmod(solver.stat().st_mode | stat.S_IEXEC)
    d = tmp_path / "in"
    d.mkdir()
    (d / "a10.tsp").write_text(_tsp(10))
    with gzip.open(d / "b20.tsp.gz", "wt") as f:
        f.write(_tsp(20))
    (d / "bad7.tsp").write_text(_tsp(7))
    (d / "big200000.tsp").write_text(_tsp(4))
    (d / "notes.txt").write_text("ignored")
    known = tmp_path / "known.txt"
    known.write_text("a10 Time: 1 Optimal Value: 15\nb20 Time: 1 Optimal Value: 30.000003\nzz Time: 1 Optimal Value: 1\n")
    out = tmp_path / "out"
    rc, times, values, errors = run_dir.run(str(d), 5, str(out), known=str(known), max_size=100000,
                                            solver=str(solver), log=lambda *a: None)
    assert set(values) == {"a10", "b20"} and set(errors) == {"bad7"}  # big200000 skipped by size
    assert open(out / "report_dir.txt").read().splitlines() == [
        "a10 Time: 0.5 Optimal Value: 15", "b20 Time: 0.5 Optimal Value: 30"]
    assert open(out / "report_dir_error.txt").read() == "bad7 ERROR: Optimal Solution not Found\n"
    rows = [ln.split() for ln in open(out / "known_answers.txt").read().splitlines()[2:]]
    assert rows[0] == ["a10", "15.0", "15.0", "0.000e+00"]
    assert rows[1][0] == "b20" and abs(float(rows[1][3]) + 1e-7) < 1e-9
    assert rows[2] == ["zz", "1.0", "missing", "-"] and rc == 1
    # the gzipped input was unpacked elsewhere: nothing written into the input directory
    assert sorted(os.listdir(d)) == ["a10.tsp", "a10.tsp.top5.F2M_report", "b20.tsp.gz", "bad7.tsp",
                                     "bad7.tsp.top5.F2M_report", "big200000.tsp", "notes.txt"]


def test_parse_report_and_names(tmp_path):
    p = tmp_path / "r"
    p.write_text("Optimal Found: Yes\nOptimal Value: 563655.760603988\nTime: 16.580322\nMAX edge rank: 18\n")
    assert run_dir.parse_report(str(p)) == ("Time: 16.580322", "Optimal Value: 563655.760603988", None)
    assert run_dir.instance_name("mona-lisa100K.tsp") == "mona-lisa100K"
    assert run_dir.instance_name("x.tsp.gz") == "x"
    assert run_dir.size_from_name("mona-lisa100K.tsp") == 100  # the reference's rule reads the first number
