"""CPU tests of the kNN restriction certificate and the warm start (oracle).

Pins: on small instances the LP over the complete graph is solved by HiGHS;
whenever the certificate counts no node, the kNN graph's optimum equals it, and
on instances where K is too small the complete graph is strictly better and the
certificate does count nodes (a certificate that always said 0 would fail).  A
warm-started solve reaches the same optimum as a cold one.
"""
import numpy as np
import pytest
from scipy.optimize import linprog
from scipy.sparse import coo_matrix

from f2m_amd import workloads as W


def complete_lp(x, y, mode, orc):
    n = len(x)
    u, w = np.triu_indices(n, 1)
    c = np.array([orc.cost(x[a], y[a], x[b], y[b], mode) for a, b in zip(u, w)])
    E = len(u)
    A = coo_matrix((np.ones(2 * E), (np.concatenate([u, w]), np.concatenate([np.arange(E)] * 2))),
                   shape=(n, E)).tocsr()
    r = linprog(c, A_eq=A, b_eq=2 * np.ones(n), bounds=(0, 1), method="highs")
    assert r.status == 0
    return float(r.fun)


@pytest.mark.parametrize("seed,mode", [(1, W.COST_NINT), (2, W.COST_RAW_FP64), (3, W.COST_NINT)])
def test_certificate_against_complete_graph(orc, seed, mode):
    x, y = W.uniform_int(60, 1000, seed)
    opt = complete_lp(x, y, mode, orc)
    saw_gap = saw_cert = False
    for K in (3, 4, 6, 10, 20):
        rp, nb, c = orc.graph_build(x, y, K, mode)
        r = orc.solve(rp, nb, c)
        st = r["stats"]
        viol = orc.knn_certify(rp, c, K, st["d_min"], r["F"])
        obj = st["objective"]
        assert obj >= opt - 1e-9 * opt
        if viol == 0:
            saw_cert = True
            assert abs(obj - opt) <= 1e-9 * opt, (K, obj, opt)
        if obj > opt + 1e-6 * opt:
            saw_gap = True
            assert viol > 0, (K, obj, opt)
    assert saw_cert and saw_gap


def test_warm_start_reaches_same_optimum(orc, golden):
    x, y, mode, K = W.c1()
    rp4, nb4, c4 = orc.graph_build(x, y, 4, mode)
    r4 = orc.solve(rp4, nb4, c4)
    rp, nb, c = orc.graph_build(x, y, K, mode)
    cold = orc.solve(rp, nb, c)
    warm = orc.solve(rp, nb, c, F_init=r4["F"])
    assert warm["stats"]["status"] == 0
    opt = [g for g in golden["lp"] if g["name"] == "C1"][0]["lp_opt"]
    assert warm["stats"]["objective"] == cold["stats"]["objective"] == opt
    # warm from the optimum of the same graph
    again = orc.solve(rp, nb, c, F_init=cold["F"])
    assert again["stats"]["objective"] == opt


def test_warm_start_converges_on_moved_points(orc):
    # a training-loop step: every instance converges (no run to the outer-round
    # cap) and the optimum equals a cold solve's
    rng = np.random.default_rng(0)
    insts, mode, K = W.c5(count=12, n=2000)
    for x, y in insts:
        x = x.astype(np.float32).astype(np.float64)
        y = y.astype(np.float32).astype(np.float64)
        rp, nb, c = orc.graph_build(x, y, K, mode)
        r = orc.solve(rp, nb, c, threads=4)
        qx = (x + 1e-4 * rng.standard_normal(len(x))).astype(np.float32).astype(np.float64)
        qy = (y + 1e-4 * rng.standard_normal(len(y))).astype(np.float32).astype(np.float64)
        rp2, nb2, c2 = orc.graph_build(qx, qy, K, mode)
        w = orc.solve(rp2, nb2, c2, threads=4, F_init=r["F"])["stats"]
        k = orc.solve(rp2, nb2, c2, threads=4)["stats"]
        assert w["status"] == 0 and w["outer_rounds"] < 50
        assert abs(w["objective"] - k["objective"]) <= 1e-9 * k["objective"]
