"""Generate the committed golden fixtures (run in the build container).

    python tests/golden/make_golden.py

Writes tests/golden/golden.json:
  * lp: LP optima from scipy.optimize.linprog(method="highs") on the kNN edge
    set of small seeded instances and of the C1 workload.  The edge set is
    built here with numpy (brute force, key (cost, id)), independently of
    the oracle, and its hash is recorded so a test can check the oracle and
    the GPU build the same graph.
  * knn_c1: numpy brute-force directed kNN of C1 (row v = 20 nearest).
  * known: the reference's published answers (examples/*/..F2M_report,
    report_dir.txt) with the LP optimum on the same edge set where computed.
The reference is read as data only (TSPLIB files, reports).
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import sys

import numpy as np
from scipy.optimize import linprog
from scipy.sparse import coo_matrix

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from f2m_amd import workloads as W  # noqa: E402


def np_cost(px, py, qx, qy, mode):
    dx = px - qx
    dy = py - qy
    c = np.sqrt(dx * dx + dy * dy)
    if mode == W.COST_NINT:
        return np.floor(c + 0.5)
    if mode == W.COST_RAW_FP32:
        return c.astype(np.float32).astype(np.float64)
    return c


def np_knn(x, y, K, mode):
    n = len(x)
    K = min(K, n - 1)
    out = np.zeros((n, K), np.int64)
    ids = np.arange(n)
    for v in range(n):
        c = np_cost(x[v], y[v], x, y, mode)
        c[v] = np.inf
        order = np.lexsort((ids, c))  # by cost, then id
        out[v] = order[:K]
    return out


def np_graph(x, y, K, mode):
    knn = np_knn(x, y, K, mode)
    n = len(x)
    adj = [dict() for _ in range(n)]
    for v in range(n):
        for w in knn[v]:
            c = float(np_cost(x[v], y[v], x[w], y[w], mode))
            adj[v][int(w)] = c
            adj[int(w)][v] = c
    rowptr = [0]
    nbr, cost = [], []
    for v in range(n):
        items = sorted(adj[v].items(), key=lambda t: (t[1], t[0]))
        nbr += [w for w, _ in items]
        cost += [c for _, c in items]
        rowptr.append(len(nbr))
    return np.array(rowptr, np.int32), np.array(nbr, np.int32), np.array(cost, np.float64), knn


def graph_hash(rowptr, nbr, cost):
    h = hashlib.sha256()
    for a in (rowptr, nbr, cost):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def lp_opt(rowptr, nbr, cost):
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    m = src < nbr
    u, w, c = src[m], nbr[m], cost[m]
    E = len(u)
    A = coo_matrix((np.ones(2 * E), (np.concatenate([u, w]), np.concatenate([np.arange(E)] * 2))),
                   shape=(n, E)).tocsr()
    r = linprog(c, A_eq=A, b_eq=2 * np.ones(n), bounds=(0, 1), method="highs")
    assert r.status == 0, r.message
    return float(r.fun), u, w, r.x


def main():
    out = {"lp": [], "known": []}
    cases = []
    for n, span, K, mode, seed in [
        (8, 30, 3, W.COST_NINT, 1), (12, 50, 4, W.COST_RAW_FP64, 2), (20, 100, 5, W.COST_RAW_FP64, 3),
        (30, 1000, 6, W.COST_RAW_FP64, 4), (50, 1000, 8, W.COST_NINT, 5), (60, 10000, 10, W.COST_RAW_FP64, 6),
        (100, 100000, 10, W.COST_RAW_FP64, 7), (200, 1000, 10, W.COST_NINT, 8), (300, 1000000, 12, W.COST_RAW_FP64, 9),
        (500, 100000, 20, W.COST_RAW_FP64, 10), (500, 300, 8, W.COST_NINT, 11), (1000, 1000000, 10, W.COST_RAW_FP32, 12),
    ]:
        x, y = W.uniform_int(n, span, seed)
        if mode == W.COST_RAW_FP32:
            x = (x / span).astype(np.float32).astype(np.float64)
            y = (y / span).astype(np.float32).astype(np.float64)
        cases.append((f"u{n}_s{span}_k{K}_m{mode}_seed{seed}", {"kind": "uniform_int", "n": n, "span": span,
                      "seed": seed, "fp32_unit": mode == W.COST_RAW_FP32}, x, y, K, mode))
    x, y, mode, K = W.c1()
    cases.append(("C1", {"kind": "c1"}, x, y, K, mode))
    for name, spec, x, y, K, mode in cases:
        rowptr, nbr, cost, knn = np_graph(x, y, K, mode)
        opt, u, w, xs = lp_opt(rowptr, nbr, cost)
        half = bool(np.all(np.isclose(xs, 0) | np.isclose(xs, 0.5) | np.isclose(xs, 1)))
        out["lp"].append({"name": name, "spec": spec, "K": K, "mode": mode, "n": int(len(x)),
                          "nslots": int(len(nbr)), "graph_sha256": graph_hash(rowptr, nbr, cost),
                          "lp_opt": opt, "lp_x_half_integral": half})
        if name == "C1":
            np.save(os.path.join(HERE, "knn_c1.npy"), knn.astype(np.int16))
        print(name, len(x), len(nbr) // 2, opt, flush=True)

    # the reference's published answers (data files of the reference)
    out["known"] = [
        {"instance": "sra104815", "K": 50, "mode": 0, "reference_opt": 246329.098557781,
         "reference_max_rank": 25, "source": "examples/vsli/sra104815.tsp.top50.F2M_report:2-4"},
        {"instance": "ara238025", "K": 20, "mode": 0, "reference_opt": 563655.760603988,
         "reference_max_rank": 18, "source": "examples/vsli/ara238025.tsp.top20.F2M_report:2-4"},
        {"instance": "ara238025", "K": 50, "mode": 0, "reference_opt": 563655.760603754,
         "reference_max_rank": 18, "source": "examples/vsli/ara238025.tsp.top50.F2M_report:2-4"},
        {"instance": "mona-lisa100K", "K": None, "mode": 0, "reference_opt": 5748948.82664499,
         "source": "report_dir.txt:4"},
    ]
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)

    # small TSPLIB instance copied as data for the CLI / reader tests
    ref = "/root/reference/examples/vsli"
    for inst in ["xqf131", "dka1376"]:
        p = os.path.join(ref, inst + ".tsp")
        if os.path.exists(p):
            with open(p, "rb") as fi, gzip.open(os.path.join(HERE, inst + ".tsp.gz"), "wb") as fo:
                fo.write(fi.read())


if __name__ == "__main__" and not {"--heavy", "--lattice", "--ara"} & set(sys.argv):
    main()


def heavy_known_answers():
    """`python tests/golden/make_golden.py --heavy`: HiGHS optimum of the LP on
    the sra104815 K=50 edge set (~1.43M edges, a few minutes), added to
    golden.json as known[...]['lp_opt_same_edges']."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from oracle import tsplib
    path = os.path.join(HERE, "golden.json")
    g = json.load(open(path))
    lab, x, y, _ = tsplib.read(os.path.join("/root/reference/examples/vsli", "sra104815.tsp"))
    rp, nb, c = O.graph_build(x, y, 50, 0)
    n = len(rp) - 1
    src = np.repeat(np.arange(n), np.diff(rp))
    m = src < nb
    u, w, cc = src[m], nb[m], c[m]
    E = len(u)
    A = coo_matrix((np.ones(2 * E), (np.concatenate([u, w]), np.concatenate([np.arange(E)] * 2))),
                   shape=(n, E)).tocsr()
    r = linprog(cc, A_eq=A, b_eq=2 * np.ones(n), bounds=(0, 1), method="highs-ds")
    assert r.status == 0, r.message
    for k in g["known"]:
        if k["instance"] == "sra104815" and k["K"] == 50:
            k["lp_opt_same_edges"] = float(r.fun)
            k["lp_method"] = "scipy.optimize.linprog(method='highs-ds')"
    json.dump(g, open(path, "w"), indent=1)
    print("sra104815 K=50 LP optimum", r.fun)


if __name__ == "__main__" and "--heavy" in sys.argv:
    heavy_known_answers()


def lattice_optima(full: bool):
    """LP optima (HiGHS) on lattice instances with many equal-cost ties (the C3
    generator): n=20000 always, n=1e6 (C3 itself) with --c3."""
    path = os.path.join(HERE, "golden.json")
    g = json.load(open(path))
    g.setdefault("lattice", [])
    sizes = [2000, 20000] + ([1_000_000] if full else [])
    for n in sizes:
        if any(e["n"] == n for e in g["lattice"]):
            continue
        x, y, mode, K = W.c3(n=n)
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        rp, nb, c = O.graph_build(x, y, K, mode)
        nn = len(rp) - 1
        src = np.repeat(np.arange(nn), np.diff(rp))
        m = src < nb
        u, w, cc = src[m], nb[m], c[m]
        E = len(u)
        A = coo_matrix((np.ones(2 * E), (np.concatenate([u, w]), np.concatenate([np.arange(E)] * 2))),
                       shape=(nn, E)).tocsr()
        r = linprog(cc, A_eq=A, b_eq=2 * np.ones(nn), bounds=(0, 1), method="highs-ipm" if n > 100000 else "highs-ds")
        assert r.status == 0, r.message
        g["lattice"].append({"n": n, "K": K, "mode": mode, "nslots": int(len(nb)),
                             "graph_sha256": graph_hash(rp, nb, c), "lp_opt": float(r.fun)})
        json.dump(g, open(path, "w"), indent=1)
        print("lattice", n, "LP optimum", r.fun, flush=True)


if __name__ == "__main__" and "--lattice" in sys.argv:
    lattice_optima("--c3" in sys.argv)


def known_answers_ara():
    """`python tests/golden/make_golden.py --ara`: the published ara238025 and
    sra104815 answers against this build's oracle, on the graphs the solver
    builds from the same files with nodes in input order (K=20 and K=50):
      * the oracle with the reference's recovery (least squares, repairs):
        its objective, MAX edge rank and x range;
      * the oracle with exact recovery: its certified objective;
      * D_all, a weak-duality lower bound of the LP over ALL pairs of points
        (tests/helpers.py: complete_graph_dual_bound) -- a published value
        below it is not the objective of any feasible x on any edge set;
      * HiGHS (interior point + crossover) on the K=20 edge set.
    Also writes tests/golden/ara238025.tsp.gz (the input, as data)."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from oracle import tsplib
    from tests.helpers import complete_graph_dual_bound
    path = os.path.join(HERE, "golden.json")
    g = json.load(open(path))
    src_dir = "/root/reference/examples/vsli"
    with open(os.path.join(src_dir, "ara238025.tsp"), "rb") as fi, \
            gzip.open(os.path.join(HERE, "ara238025.tsp.gz"), "wb", compresslevel=9) as fo:
        fo.write(fi.read())
    g["known"] = [k for k in g["known"] if k["instance"] not in ("ara238025",)]
    for inst, Ks in (("ara238025", (20, 50)), ("sra104815", (50,))):
        lab, x, y, _ = tsplib.read(os.path.join(src_dir, inst + ".tsp"))
        for K in Ks:
            rep = open(os.path.join(src_dir, f"{inst}.tsp.top{K}.F2M_report")).read().splitlines()
            ent = [k for k in g["known"] if k["instance"] == inst and k["K"] == K]
            e = ent[0] if ent else {"instance": inst, "K": K, "mode": 0,
                                    "source": f"examples/vsli/{inst}.tsp.top{K}.F2M_report:2-4"}
            e["reference_opt"] = float(rep[1].split(": ")[1])
            e["reference_max_rank"] = int(rep[3].split(": ")[1])
            rp, nb, c = O.graph_build(x, y, K, 0)
            r = O.solve(rp, nb, c)
            r0 = O.solve(rp, nb, c, O.default_opts(exact=0))
            Dall, nviol = complete_graph_dual_bound(x, y, rp, nb, c, r["F"], r["stats"]["d_min"], K, 0)
            e.update({"node_order": "input", "nslots": int(len(nb)), "graph_sha256": graph_hash(rp, nb, c),
                      "oracle_ref_recovery_opt": r0["stats"]["objective"],
                      "oracle_ref_recovery_max_rank": r0["stats"]["max_rank"],
                      "oracle_ref_recovery_repairs": r0["stats"]["n_repairs"],
                      "oracle_ref_recovery_x_range": [float(r0["x"].min()), float(r0["x"].max())],
                      "oracle_exact_opt": r["stats"]["objective"], "oracle_exact_gap": r["stats"]["exact_gap"],
                      "oracle_exact_max_rank": r["stats"]["max_rank"],
                      "dual_bound_all_pairs": Dall, "dual_bound_endpoints_checked": nviol})
            if inst == "ara238025" and K == 20:
                n = len(rp) - 1
                s_ = np.repeat(np.arange(n), np.diff(rp))
                m = s_ < nb
                u, w, cc = s_[m], nb[m], c[m]
                E = len(u)
                A = coo_matrix((np.ones(2 * E), (np.concatenate([u, w]), np.concatenate([np.arange(E)] * 2))),
                               shape=(n, E)).tocsr()
                lp = linprog(cc, A_eq=A, b_eq=2 * np.ones(n), bounds=(0, 1), method="highs-ipm")
                assert lp.status == 0, lp.message
                e["lp_opt_same_edges"] = float(lp.fun)
                e["lp_method"] = "scipy.optimize.linprog(method='highs-ipm')"
            if not ent:
                g["known"].append(e)
            print(inst, K, {k: e[k] for k in e if k not in ("graph_sha256",)}, flush=True)
            json.dump(g, open(path, "w"), indent=1)


if __name__ == "__main__" and "--ara" in sys.argv:
    known_answers_ara()
