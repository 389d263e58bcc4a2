"""GPU tests of the CLI (counterpart of RUN.cpp), of the known-answer runner
(counterpart of run_dir.py) and of the reference's published known answers
(sra104815 K=50, ara238025 K=20 and K=50)."""
import gzip
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from tests.conftest import say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "f2m_amd", "f2m_solver")


def _unpack(tmp, inst):
    p = os.path.join(tmp, inst + ".tsp")
    with gzip.open(os.path.join(ROOT, "tests", "golden", inst + ".tsp.gz"), "rb") as fi, open(p, "wb") as fo:
        fo.write(fi.read())
    return p


@pytest.mark.parametrize("inst,K,order", [("xqf131", 10, "morton"), ("dka1376", 20, "morton"),
                                          ("dka1376", 20, "input")])
def test_cli_files_and_optimum(orc, inst, K, order):
    from oracle import tsplib
    from tests.helpers import morton_order
    tmp = tempfile.mkdtemp()
    try:
        p = _unpack(tmp, inst)
        say("cli", inst, K, order)
        flags = ["--quiet"] + (["--input-order"] if order == "input" else [])
        r = subprocess.run([CLI, p, str(K)] + flags, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        rep = open(p + f".top{K}.F2M_report").read().splitlines()
        assert rep[0] == "Optimal Found: Yes"
        m = re.match(r"Optimal Value: (\S+)$", rep[1])
        assert m and rep[2].startswith("Time: ") and rep[3].startswith("MAX edge rank: ")
        obj = float(m.group(1))
        # same instance through the oracle (parse restatement + kNN + solve),
        # nodes numbered as the CLI numbers them
        lab, x, y, _ = tsplib.read(p)
        perm = morton_order(x, y) if order == "morton" else np.arange(len(x))
        x, y, lab = x[perm], y[perm], lab[perm]
        rp, nb, c = orc.graph_build(x, y, K, 0)
        ro = orc.solve(rp, nb, c)
        assert abs(obj - ro["stats"]["objective"]) <= 1e-9 * obj + 1e-9
        assert int(rep[3].split(": ")[1]) == ro["stats"]["max_rank"]
        # solution file: labels, degree 2 at every node
        deg = {}
        for line in open(p + f".top{K}.F2M_solution"):
            a, b, xv = line.split()
            assert int(a) < int(b)
            deg[int(a)] = deg.get(int(a), 0) + float(xv)
            deg[int(b)] = deg.get(int(b), 0) + float(xv)
        assert set(deg) == set(lab.tolist())
        assert max(abs(v - 2) for v in deg.values()) < 1e-3
        # .DM: one line per kept edge, 0-based input indices
        dm = [ln.split() for ln in open(p + f".top{K}.DM").read().splitlines()]
        assert len(dm) == len(nb) // 2
        src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        mk = src < nb
        want = {(min(perm[a], perm[b]), max(perm[a], perm[b])) for a, b in zip(src[mk], nb[mk])}
        assert {(int(a), int(b)) for a, b, _ in dm} == want
        # usage message and exit code of the reference (exit(0))
        r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "usage" in r.stdout
    finally:
        shutil.rmtree(tmp)


def test_cli_adaptive_k(orc):
    from oracle import tsplib
    from tests.helpers import morton_order
    tmp = tempfile.mkdtemp()
    try:
        p = _unpack(tmp, "xqf131")
        say("cli adaptive xqf131 from K=3")
        r = subprocess.run([CLI, p, "3", "--adaptive"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        rep = open(p + ".top3.F2M_report").read().splitlines()
        assert rep[0] == "Optimal Found: Yes" and rep[4].startswith("kNN K: ") and rep[4].endswith("(certified)")
        Kf = int(rep[4].split()[2])
        assert Kf > 3 and "nodes failing the kNN certificate 0" in r.stdout
        lab, x, y, _ = tsplib.read(p)
        pm = morton_order(x, y)
        rp, nb, c = orc.graph_build(x[pm], y[pm], Kf, 0)
        ro = orc.solve(rp, nb, c)
        assert abs(float(rep[1].split(": ")[1]) - ro["stats"]["objective"]) <= 1e-9 * ro["stats"]["objective"]
        assert orc.knn_certify(rp, c, Kf, ro["stats"]["d_min"], ro["F"]) == 0
    finally:
        shutil.rmtree(tmp)


def test_known_answer_sra104815_k50(ctx, orc, golden):
    from f2m_amd.solver import default_opts, read_tsplib
    known = [k for k in golden["known"] if k["instance"] == "sra104815"][0]
    tmp = tempfile.mkdtemp()
    try:
        p = _unpack(tmp, "sra104815")
        pts = read_tsplib(p)
        say("sra104815 K=50 solve, reference recovery")
        g = ctx.knn_build(pts.x, pts.y, 50, 0)
        rp, nb, c = ctx.graph_csr(g)
        st = ctx.solve(g, default_opts(exact=0))
        res = ctx.result()
        r = orc.solve(rp, nb, c, orc.default_opts(exact=0))
        # the GPU follows the oracle exactly, repairs included
        assert np.array_equal(res["F"], r["F"]) and np.array_equal(res["comb_root"], r["comb_root"])
        assert st["n_repairs"] == r["stats"]["n_repairs"] and st["status"] == r["stats"]["status"] == 0
        assert abs(st["objective"] - r["stats"]["objective"]) <= 1e-9 * r["stats"]["objective"]
        assert st["max_rank"] == known["reference_max_rank"] == 25
        # the reference's published value is a least-squares artefact as well
        # (x outside [0,1]); ours lies within 2e-6 relative of it
        assert abs(st["objective"] - known["reference_opt"]) <= 2e-6 * known["reference_opt"]
        assert res["x"].min() < 0 or res["x"].max() > 1
        # default: exact recovery, the LP optimum of HiGHS on the same edges
        say("sra104815 K=50 solve, exact recovery")
        st = ctx.solve(g)
        res = ctx.result()
        r = orc.solve(rp, nb, c)
        for k in ("exact_ran", "exact_phases", "exact_rounds", "exact_certified", "exact_gap", "status"):
            assert st[k] == r["stats"][k], k
        assert np.array_equal(res["x"], r["x"]) and np.array_equal(res["F"], r["F"])
        opt = known["lp_opt_same_edges"]
        assert st["exact_certified"] == 1 and abs(st["objective"] - opt) <= 1e-9 * opt
        assert res["x"].min() == 0.0 and res["x"].max() == 1.0
        # the published value lies below a lower bound of the LP over all pairs
        # of points (weak duality, computed by make_golden.py --ara): no
        # feasible x on any edge set reaches it
        assert known["reference_opt"] < known["dual_bound_all_pairs"] <= st["objective"]
    finally:
        shutil.rmtree(tmp)


@pytest.mark.parametrize("K", [20, 50])
def test_known_answer_ara238025(ctx, golden, K):
    """The reference's published ara238025 answers (examples/vsli/
    ara238025.tsp.top{20,50}.F2M_report:2-4), nodes in input order."""
    from f2m_amd.solver import default_opts, read_tsplib
    from tests.helpers import certificate, graph_hash
    known = [k for k in golden["known"] if k["instance"] == "ara238025" and k["K"] == K][0]
    tmp = tempfile.mkdtemp()
    try:
        p = _unpack(tmp, "ara238025")
        pts = read_tsplib(p)
        assert len(pts) == 238025 and pts.n_duplicates == 0
        say("ara238025 K", K, "kNN graph")
        g = ctx.knn_build(pts.x, pts.y, K, 0)
        rp, nb, c = ctx.graph_csr(g)
        assert graph_hash(rp, nb, c) == known["graph_sha256"]  # the oracle's graph
        say("ara238025 K", K, "solve, reference recovery")
        st = ctx.solve(g, default_opts(exact=0))
        res = ctx.result()
        assert st["status"] == 0 and st["n_repairs"] == known["oracle_ref_recovery_repairs"]
        assert abs(st["objective"] - known["oracle_ref_recovery_opt"]) <= 1e-9 * known["oracle_ref_recovery_opt"]
        # the published value, and its MAX edge rank
        assert abs(st["objective"] - known["reference_opt"]) <= 1e-9 * known["reference_opt"]
        assert st["max_rank"] == known["reference_max_rank"] == 18
        # ... which comes from a least-squares x outside [0, 1]
        assert res["x"].min() < 0 and res["x"].max() > 1
        say("ara238025 K", K, "solve, exact recovery")
        st = ctx.solve(g)
        res = ctx.result()
        assert st["status"] == 0 and st["exact_certified"] == 1
        assert abs(st["objective"] - known["oracle_exact_opt"]) <= 1e-12 * known["oracle_exact_opt"]
        if "lp_opt_same_edges" in known:  # HiGHS on the same edges
            assert abs(st["objective"] - known["lp_opt_same_edges"]) <= 1e-9 * known["lp_opt_same_edges"]
        assert res["x"].min() == 0.0 and res["x"].max() == 1.0
        P, D = certificate(rp, nb, c, res["x"], res["F"], st["d_min"])
        assert abs(P - st["objective"]) <= 1e-9 * P and 0 <= P - D <= 1e-9 * P
        # the published value lies below a lower bound of the LP over all
        # pairs of points: it is not the objective of any feasible x
        assert known["reference_opt"] < known["dual_bound_all_pairs"] <= P
        if K == 20:  # the same through the CLI and its files
            say("ara238025 CLI --input-order --reference-recovery")
            r = subprocess.run([CLI, p, "20", "--quiet", "--input-order", "--reference-recovery"], capture_output=True,
                               text=True, timeout=300)
            assert r.returncode == 0, r.stderr
            rep = open(p + ".top20.F2M_report").read().splitlines()
            assert rep[0] == "Optimal Found: Yes" and rep[3] == "MAX edge rank: 18"
            v = float(rep[1].split(": ")[1])
            assert abs(v - known["reference_opt"]) <= 1e-9 * known["reference_opt"]
    finally:
        shutil.rmtree(tmp)


def test_run_dir_known_answers(orc):
    """The known-answer runner (counterpart of run_dir.py) over a directory of
    gzipped TSPLIB files: tables in the reference's format and the
    differences to a known-answer file holding the oracle's optima."""
    from f2m_amd.cli import run_dir
    from oracle import tsplib
    from tests.helpers import morton_order
    tmp = tempfile.mkdtemp()
    try:
        src = os.path.join(tmp, "in")
        os.makedirs(src)
        known = []
        for inst in ("xqf131", "dka1376"):
            shutil.copy(os.path.join(ROOT, "tests", "golden", inst + ".tsp.gz"), src)
            p = _unpack(tmp, inst)
            lab, x, y, _ = tsplib.read(p)
            pm = morton_order(x, y)
            rp, nb, c = orc.graph_build(x[pm], y[pm], 10, 0)
            known.append((inst, orc.solve(rp, nb, c)["stats"]["objective"]))
        kf = os.path.join(tmp, "known.txt")
        with open(kf, "w") as f:
            for inst, v in known:
                f.write(f"{inst} Time: 0 Optimal Value: {v!r}\n")
        out = os.path.join(tmp, "out")
        say("run_dir K=10")
        rc, times, values, errors = run_dir.run(src, 10, out, known=kf, log=lambda *a: None)
        assert rc == 0 and not errors and set(values) == {"xqf131", "dka1376"}
        lines = open(os.path.join(out, "report_dir.txt")).read().splitlines()
        assert len(lines) == 2 and all(re.match(r"\S+ Time: \S+ Optimal Value: \S+$", ln) for ln in lines)
        assert open(os.path.join(out, "report_dir_error.txt")).read() == ""
        rows = [ln.split() for ln in open(os.path.join(out, "known_answers.txt")).read().splitlines()[2:]]
        assert len(rows) == 2 and all(abs(float(r[3])) <= 1e-9 for r in rows)
    finally:
        shutil.rmtree(tmp)
