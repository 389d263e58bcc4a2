"""CPU tests pinning the oracle (the checker) before it is trusted.

Each pin is chosen so that a plausible mistake in the restatement fails:
  * kNN: equality with an independent numpy brute force (lexsort by cost, id)
    and grid == brute force on tie-heavy inputs;
  * graph: the symmetrised CSR hashes equal the fixtures built by numpy;
  * optimum: HiGHS LP optima (scipy) on fixed edge sets, within 1e-9 rel.;
  * sweep: closed forms on hand cases (a dropped term, wrong order statistic
    or wrong sign changes the value);
  * partition: invariants of a Hoare partition + the EF raise rule;
  * least squares: the minimum-norm solution (numpy pinv) of consistent systems;
  * backward: central finite differences of the closed form.
"""
import math
import os

import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.helpers import feasibility, graph_hash, instance_from_spec, sym_check

HERE = os.path.dirname(os.path.abspath(__file__))


def test_knn_c1_matches_numpy_bruteforce(orc):
    x, y, mode, K = W.c1()
    ref = np.load(os.path.join(HERE, "golden", "knn_c1.npy")).astype(np.int32)
    assert np.array_equal(orc.knn_directed(x, y, K, mode, brute=True), ref)
    assert np.array_equal(orc.knn_directed(x, y, K, mode, brute=False), ref)


@pytest.mark.parametrize("n,span,K,mode,seed", [
    (2000, 40, 10, W.COST_NINT, 1),       # heavy ties
    (3000, 10**6, 20, W.COST_RAW_FP64, 2),
    (1500, 10**5, 16, W.COST_RAW_FP32, 3),
    (50, 10, 49, W.COST_NINT, 4),         # K = n - 1 after dedup clamp
])
def test_knn_grid_equals_bruteforce(orc, n, span, K, mode, seed):
    x, y = W.uniform_int(n, span, seed)
    a = orc.knn_directed(x, y, K, mode, brute=True)
    b = orc.knn_directed(x, y, K, mode, brute=False)
    assert np.array_equal(a, b)


def test_knn_line_and_clusters(orc):
    # degenerate geometry: collinear points and a tight cluster far away
    x = np.concatenate([np.arange(300, dtype=float), 1e6 + np.arange(20) * 0.5])
    y = np.concatenate([np.zeros(300), 1e6 + (np.arange(20) % 3) * 0.25])
    for mode in (W.COST_RAW_FP64, W.COST_NINT):
        assert np.array_equal(orc.knn_directed(x, y, 8, mode, brute=True),
                              orc.knn_directed(x, y, 8, mode, brute=False))


def test_graph_matches_fixture_hash_and_lp_optimum(orc, golden):
    for case in golden["lp"]:
        x, y = instance_from_spec(case["spec"])
        rp, nb, c = orc.graph_build(x, y, case["K"], case["mode"])
        assert graph_hash(rp, nb, c) == case["graph_sha256"], case["name"]
        r = orc.solve(rp, nb, c)
        st = r["stats"]
        assert st["status"] == 0, case["name"]
        assert abs(st["objective"] - case["lp_opt"]) <= 1e-9 * abs(case["lp_opt"]), (case["name"], st)
        # the dual value is a lower bound and meets the optimum: to 1e-7 when
        # the reference recovery succeeded; when its x was infeasible and the
        # exact recovery ran, within the certified gap (< 1/2 on integer costs)
        assert st["dual"] <= case["lp_opt"] * (1 + 1e-12) + 1e-9
        if st["exact_ran"]:
            assert st["exact_certified"] == 1 and set(np.unique(r["x"])) <= {0.0, 0.5, 1.0}, case["name"]
            tol = 0.5 if case["mode"] == W.COST_NINT else 1e-10 * abs(case["lp_opt"])
            assert case["lp_opt"] - st["dual"] <= tol, (case["name"], st)
        else:
            assert abs(st["dual"] - case["lp_opt"]) <= 1e-7 * abs(case["lp_opt"]), (case["name"], st)
        err, lo, hi = feasibility(rp, nb, r["x"])
        assert err < 1e-6 and lo >= -1e-6 and hi <= 1 + 1e-6, case["name"]
        assert sym_check(rp, nb, r["x"])


def test_sweep_closed_forms(orc):
    # star: node 0 joined to 1, 2, 3 with D = 1, 2, 3; F = 0
    rowptr = np.array([0, 3, 4, 5, 6], np.int32)
    nbr = np.array([1, 2, 3, 0, 0, 0], np.int32)
    D = np.array([1.0, 2.0, 3.0, 1.0, 2.0, 3.0])
    length = np.diff(rowptr).astype(np.int32)
    F0 = np.zeros(4)
    F1, g = orc.sweep(rowptr, nbr, D, length, F0)
    # node 0: t = (1,2,3), g = -(2+3)/2; leaves: one edge + two 1e6 sentinels
    assert g[0] == -2.5
    assert g[1] == -(1e6 + 1e6) / 2
    assert np.array_equal(F1, F0 + g)
    # d uses half of each endpoint potential: F = (2, 4, 0, 0) at node 0
    F = np.array([2.0, 4.0, 0.0, 0.0])
    _, g = orc.sweep(rowptr, nbr, D, length, F)
    # d = 1 + 1 + 2 = 4, 2 + 1 = 3, 3 + 1 = 4  -> sorted (3, 4, 4)
    assert g[0] == -(4 + 4) / 2
    # a list shorter than its degree only sees its prefix
    _, g = orc.sweep(rowptr, nbr, D, np.array([2, 1, 1, 1], np.int32), F0)
    assert g[0] == -(2 + 1e6) / 2


def test_sweep_is_jacobi(orc):
    # all nodes read the same old F: a triangle with equal costs stays symmetric
    rowptr = np.array([0, 2, 4, 6], np.int32)
    nbr = np.array([1, 2, 0, 2, 0, 1], np.int32)
    D = np.ones(6)
    F = np.zeros(3)
    for _ in range(5):
        F, g = orc.sweep(rowptr, nbr, D, np.full(3, 2, np.int32), F)
    assert F[0] == F[1] == F[2]


def test_partition_semantics(orc):
    rng = np.random.default_rng(0)
    for trial in range(200):
        size = int(rng.integers(0, 30))
        d = rng.normal(size=size).round(1)
        thr = float(rng.normal())
        p, dd, idx = orc.qpart(d, thr, raise_=False)
        assert sorted(dd.tolist()) == sorted(d.tolist())
        assert np.array_equal(dd, d[idx])
        if size >= 2:
            assert p == int((d <= thr).sum())
            assert np.all(dd[:p] <= thr) and np.all(dd[p:] > thr)
        if size >= 3:
            p2, dd2, _ = orc.qpart(d, thr, raise_=True)
            t = thr
            while (d <= t).sum() < 3:
                t += 1.0
            assert p2 == int((d <= t).sum()) and p2 >= 3
    # single-element list: the reference's loop never runs and reports 1
    assert orc.qpart([5.0], 0.0, raise_=False)[0] == 1


def test_least_squares_is_min_norm(orc):
    rng = np.random.default_rng(1)
    for trial in range(20):
        nv = int(rng.integers(3, 12))
        m = int(rng.integers(2, nv))
        rows = [sorted(set(rng.integers(0, nv, size=int(rng.integers(1, 4))).tolist())) for _ in range(m)]
        A = np.zeros((m, nv))
        for j, r in enumerate(rows):
            A[j, r] = 1
        xt = rng.integers(0, 2, size=nv).astype(float)
        b = (A @ xt).astype(int)
        go, x, err = orc.axb(rows, b, nv)
        if go == 0:
            continue
        assert err < 1e-8
        assert np.allclose(A @ x, b, atol=1e-6)
        assert np.allclose(x, np.linalg.pinv(A) @ b, atol=1e-5)


def test_backward_closed_form_vs_fd(orc):
    rng = np.random.default_rng(2)
    n = 40
    px, py = rng.random(n), rng.random(n)
    rp, nb, c = orc.graph_build(px, py, 6, W.COST_RAW_FP64)
    x = np.zeros(len(nb))
    src = np.repeat(np.arange(n), np.diff(rp))
    w = rng.random(len(nb))
    key = np.minimum(src, nb) * n + np.maximum(src, nb)
    _, inv = np.unique(key, return_inverse=True)
    x = w[np.unique(key, return_index=True)[1]][inv]  # same value on both slots

    def V(px_, py_):
        m = src < nb
        return float(np.sum(x[m] * np.hypot(px_[src[m]] - px_[nb[m]], py_[src[m]] - py_[nb[m]])))

    gx, gy = orc.bound_backward(rp, nb, x, px, py)
    h = 1e-6
    for u in range(0, n, 7):
        e = np.zeros(n)
        e[u] = h
        fdx = (V(px + e, py) - V(px - e, py)) / (2 * h)
        fdy = (V(px, py + e) - V(px, py - e)) / (2 * h)
        assert math.isclose(gx[u], fdx, rel_tol=1e-6, abs_tol=1e-8)
        assert math.isclose(gy[u], fdy, rel_tol=1e-6, abs_tol=1e-8)


def test_schedule_counts_are_deterministic(orc):
    x, y, mode, K = W.c1()
    rp, nb, c = orc.graph_build(x, y, K, mode)
    a = orc.solve(rp, nb, c, threads=1)
    b = orc.solve(rp, nb, c, threads=4)
    assert a["stats"] == b["stats"]
    assert np.array_equal(a["F"], b["F"]) and np.array_equal(a["x"], b["x"])
    assert np.array_equal(a["comb_root"], b["comb_root"])


def test_tsplib_reader_restatement():
    from oracle import tsplib
    text = ("NAME : t\nCOMMENT : c\nNODE_COORD_SECTION\r\n1 0.0004 1.0\r\n2 3.00051 4\r\n3 0.0 1.0\r\n"
            "4 -2.5005 7\nEOF\n9 9 9\n")
    lab, x, y, d = tsplib.read(text, is_text=True)
    assert lab.tolist() == [1, 2, 4]
    assert x.tolist() == [0.0, 3.001, -2.501] and y.tolist() == [1.0, 4.0, 7.0]
    assert d == {1: [3]}
