"""CPU tests pinning the oracle's exact primal recovery (f2m_exact_oracle.cpp).

The stage has no counterpart in the reference.  Its pins:
  * LP optima of HiGHS on fixed edge sets (tests/golden/golden.json: the
    uniform cases, C1 and tie-heavy lattices), from warm and cold potentials;
  * feasibility: x in {0, 1/2, 1}, degree exactly 2, equal on both slots;
  * the certificate recomputed in numpy from the returned x and potentials:
    P equals the objective, D <= the HiGHS optimum (weak duality), and
    P - D equals the reported gap (a wrong price-to-dual map, sign or
    orientation of the double cover breaks one of these);
  * error paths: asymmetric CSR, a node of degree 1, an infeasible LP.
"""
import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.helpers import certificate, feasibility, instance_from_spec, sym_check


def _check_solution(rp, nb, c, r, opt, integral):
    st = r["stats"]
    x = r["x"]
    err, lo, hi = feasibility(rp, nb, x)
    assert err == 0.0 and lo >= 0.0 and hi <= 1.0
    assert set(np.unique(x)) <= {0.0, 0.5, 1.0}
    assert sym_check(rp, nb, x)
    d_min = float(c[c > 0].min())
    P, D = certificate(rp, nb, c, x, r["F"], d_min)
    assert D <= opt + 1e-9 * abs(opt) + 1e-9
    tol = 0.5 if integral else 1e-10 * abs(opt)
    assert P - D <= tol
    assert abs((P - D) - st["exact_gap"]) <= 1e-7 * max(1.0, abs(opt))
    assert abs(P - opt) <= (0.0 if integral else 1e-9 * abs(opt)), (P, opt)


def test_exact_lattice_matches_highs(orc, golden):
    e = [e for e in golden["lattice"] if e["n"] == 2000][0]
    x, y, mode, K = W.c3(n=e["n"])
    rp, nb, c = orc.graph_build(x, y, K, mode)
    # the reference recovery fails on this tie-heavy lattice ...
    ref = orc.solve(rp, nb, c, orc.default_opts(exact=0))
    err, lo, hi = feasibility(rp, nb, ref["x"])
    assert ref["stats"]["status"] != 0 or lo < -1e-6 or hi > 1 + 1e-6 or err > 1e-6
    # ... and the default (exact when it fails) gives the LP optimum exactly
    r = orc.solve(rp, nb, c)
    st = r["stats"]
    assert st["exact_ran"] == 1 and st["exact_certified"] == 1 and st["status"] == 0
    assert st["objective"] == e["lp_opt"]
    _check_solution(rp, nb, c, r, e["lp_opt"], True)
    # comb_root: components of the x = 1/2 edges, rooted at their smallest id
    half = np.repeat(np.arange(len(rp) - 1), np.diff(rp))[r["x"] == 0.5]
    assert set(np.unique(half).tolist()) == set(np.nonzero(r["comb_root"] >= 0)[0].tolist())
    roots = r["comb_root"][r["comb_root"] >= 0]
    assert np.all(roots <= np.nonzero(r["comb_root"] >= 0)[0])
    assert st["n_combs"] == len(np.unique(roots))


def test_exact_cold_and_warm_start_on_golden(orc, golden):
    for case in golden["lp"]:
        x, y = instance_from_spec(case["spec"])
        rp, nb, c = orc.graph_build(x, y, case["K"], case["mode"])
        integral = case["mode"] == W.COST_NINT
        n = len(rp) - 1
        rc, r = orc.exact(rp, nb, c, np.zeros(n))
        assert rc == 0, case["name"]
        _check_solution(rp, nb, c, r, case["lp_opt"], integral)
        # forced exact after the reference pipeline (warm potentials)
        w = orc.solve(rp, nb, c, orc.default_opts(exact=2))
        assert w["stats"]["exact_ran"] == 1 and w["stats"]["status"] == 0, case["name"]
        _check_solution(rp, nb, c, w, case["lp_opt"], integral)


def test_exact_matches_reference_where_it_succeeds(orc, golden):
    # where the reference recovery gives a feasible x the default leaves it alone
    case = [c for c in golden["lp"] if c["name"].startswith("u100_")][0]
    x, y = instance_from_spec(case["spec"])
    rp, nb, c = orc.graph_build(x, y, case["K"], case["mode"])
    a = orc.solve(rp, nb, c)
    b = orc.solve(rp, nb, c, orc.default_opts(exact=0))
    assert a["stats"]["exact_ran"] == 0
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["F"], b["F"])
    # C1: the reference's least squares leaves x outside [0,1]; exact fixes it
    x, y, mode, K = W.c1()
    rp, nb, c = orc.graph_build(x, y, K, mode)
    b = orc.solve(rp, nb, c, orc.default_opts(exact=0))
    assert b["x"].min() < -1e-3
    a = orc.solve(rp, nb, c)
    assert a["stats"]["exact_ran"] == 1 and a["x"].min() == 0.0


def _csr(n, edges):
    adj = [[] for _ in range(n)]
    for u, w, cst in edges:
        adj[u].append((cst, w))
        adj[w].append((cst, u))
    rp = [0]
    nb, cc = [], []
    for v in range(n):
        for cst, w in sorted(adj[v]):
            nb.append(w)
            cc.append(cst)
        rp.append(len(nb))
    return np.array(rp, np.int32), np.array(nb, np.int32), np.array(cc, np.float64)


def test_exact_error_paths(orc):
    # bowtie: two triangles sharing node 0; nodes 1-4 force x = 1 on all six
    # edges, so node 0 would carry 4: no fractional 2-matching exists
    rp, nb, c = _csr(5, [(0, 1, 1.0), (0, 2, 2.0), (1, 2, 1.0), (0, 3, 1.0), (0, 4, 3.0), (3, 4, 1.0)])
    rc, r = orc.exact(rp, nb, c, np.zeros(5), opts=orc.default_opts(exact_max_rounds=500))
    assert rc == 1 and r["stats"]["exact_certified"] == 0 and r["stats"]["exact_rounds"] == 500
    # a node of degree 1
    rp, nb, c = _csr(4, [(0, 1, 1.0), (1, 2, 1.0), (0, 2, 1.0), (2, 3, 1.0)])
    assert orc.exact(rp, nb, c, np.zeros(4))[0] == -2
    # asymmetric costs
    rp, nb, c = _csr(3, [(0, 1, 1.0), (1, 2, 1.0), (0, 2, 1.0)])
    c = c.copy()
    c[0] = 5.0
    assert orc.exact(rp, nb, c, np.zeros(3))[0] == -1
    # a triangle: x = 1 on every edge
    rp, nb, c = _csr(3, [(0, 1, 2.0), (1, 2, 3.0), (0, 2, 4.0)])
    rc, r = orc.exact(rp, nb, c, np.zeros(3))
    assert rc == 0 and np.all(r["x"] == 1.0)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_exact_tiny_bruteforce(orc, seed):
    """Tiny graphs: the exact optimum equals the minimum over all half-integral
    x in {0, 1/2, 1}^E with degree 2 (the LP has a half-integral optimum)."""
    import itertools
    rng = np.random.default_rng(seed)
    n = 6
    pts = rng.integers(0, 20, size=(n, 2)).astype(float)
    edges = []
    for u in range(n):
        for w in range(u + 1, n):
            if rng.random() < 0.7:
                edges.append((u, w, float(np.floor(np.hypot(*(pts[u] - pts[w])) + 0.5)) + 1.0))
    edges = edges[:11]
    rp, nb, c = _csr(n, edges)
    if np.any(np.diff(rp) < 2):
        pytest.skip("degree below 2")
    best = None
    for xs in itertools.product((0.0, 0.5, 1.0), repeat=len(edges)):
        deg = np.zeros(n)
        for (u, w, _), xv in zip(edges, xs):
            deg[u] += xv
            deg[w] += xv
        if np.all(deg == 2):
            val = sum(cc * xv for (_, _, cc), xv in zip(edges, xs))
            best = val if best is None else min(best, val)
    rc, r = orc.exact(rp, nb, c, np.zeros(n), opts=orc.default_opts(exact_max_rounds=20000))
    if best is None:
        assert rc != 0
    else:
        assert rc == 0
        _check_solution(rp, nb, c, r, best, True)
