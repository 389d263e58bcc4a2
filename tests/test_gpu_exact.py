"""GPU tests of the exact primal recovery (f2m_exact.hip).

Bar: bit for bit equal to the oracle's restatement (potentials, x,
components, phase and round counts, certified gap), on warm starts after the
reference pipeline and on forced runs; the objective equal to the HiGHS LP
optimum (tests/golden/golden.json) exactly on integer costs, including the
full C3 lattice (n = 1e6) where only the GPU runs; x in {0, 1/2, 1} with
degree exactly 2; the error paths of the oracle.
"""
import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.conftest import say
from tests.helpers import certificate, feasibility, sym_check

pytestmark = pytest.mark.gpu

EXACT_KEYS = ("sweeps", "status", "n_combs", "n_irregular_a", "max_rank", "exact_ran", "exact_phases",
              "exact_rounds", "exact_certified", "exact_gap")


def _compare(ctx, orc, rp, nb, c, name, **kw):
    from f2m_amd.solver import default_opts
    g = ctx.graph_from_csr(rp, nb, c)
    say("exact solve", name, len(rp) - 1, kw)
    st = ctx.solve(g, default_opts(**kw))
    res = ctx.result()
    r = orc.solve(rp, nb, c, orc.default_opts(**kw))
    so = r["stats"]
    for k in EXACT_KEYS:
        assert st[k] == so[k], (name, k, st[k], so[k])
    assert np.array_equal(res["F"], r["F"]), name
    assert np.array_equal(res["x"], r["x"]), name
    assert np.array_equal(res["comb_root"], r["comb_root"]), name
    assert abs(st["objective"] - so["objective"]) <= 1e-12 * abs(so["objective"])
    return st, res


def _feasible(rp, nb, x):
    err, lo, hi = feasibility(rp, nb, x)
    assert err == 0.0 and lo >= 0.0 and hi <= 1.0
    assert set(np.unique(x)) <= {0.0, 0.5, 1.0}
    assert sym_check(rp, nb, x)


@pytest.mark.parametrize("n", [2000, 20000])
def test_exact_lattice_bit_exact_and_optimal(ctx, orc, golden, n):
    e = [e for e in golden["lattice"] if e["n"] == n][0]
    x, y, mode, K = W.c3(n=n)
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    st, res = _compare(ctx, orc, rp, nb, c, f"lattice{n}")
    assert st["exact_ran"] == 1 and st["exact_certified"] == 1 and st["status"] == 0
    assert st["objective"] == e["lp_opt"]
    _feasible(rp, nb, res["x"])


EXACT_ENV = ("F2M_EXACT_WG", "F2M_EXACT_REG", "F2M_EXACT_FUSE")


def _variants_equal(ctx, orc, monkeypatch, rp, nb, c, name, variants, **kw):
    ref = None
    for env in variants:
        for k in EXACT_ENV:
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        st, res = _compare(ctx, orc, rp, nb, c, f"{name} {env}", **kw)
        assert st["exact_certified"] == 1
        if ref is None:
            ref = (st, res)
        else:
            for k in EXACT_KEYS:
                assert st[k] == ref[0][k], (env, k)
            assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    for k in EXACT_ENV:
        monkeypatch.delenv(k, raising=False)


def test_exact_rounds_in_workgroup_or_grid(ctx, orc, monkeypatch):
    # the rounds of small active lists run inside one workgroup (x_rounds_wg),
    # the others as grid-wide steps: all rounds in grid steps (WG 0), the
    # defaults, all rounds inside one workgroup, discharges gathering their
    # row per push (REG 0) and a wait after every launch group (FUSE 0) give
    # the same bits
    x, y, mode, K = W.c3(n=20000)
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    _variants_equal(ctx, orc, monkeypatch, rp, nb, c, "lattice20000",
                    [{}, {"F2M_EXACT_WG": "0"}, {"F2M_EXACT_WG": "100000000"}, {"F2M_EXACT_REG": "0"},
                     {"F2M_EXACT_FUSE": "0"}])


def test_exact_rows_beyond_register_cache(ctx, orc, monkeypatch):
    # rows of more than 64 slots discharge by gathers, shorter ones from
    # registers; K = 56 gives both in one graph (forced exact stage)
    x, y = W.uniform_int(3000, 10**6, 23)
    g = ctx.knn_build(x, y, 56, W.COST_NINT)
    rp, nb, c = ctx.graph_csr(g)
    deg = np.diff(rp)
    assert deg.min() <= 64 < deg.max()
    _variants_equal(ctx, orc, monkeypatch, rp, nb, c, "uniform3000 K56",
                    [{}, {"F2M_EXACT_REG": "0"}, {"F2M_EXACT_WG": "0", "F2M_EXACT_FUSE": "0"}], exact=2)


def test_exact_forced_raw_costs(ctx, orc):
    # non-integer costs: certified to 1e-10 relative; forced after a pipeline
    # whose own recovery succeeded, so the optimum must not move
    x, y = W.uniform_int(20000, 10**7, 21)
    g = ctx.knn_build(x, y, 10, W.COST_RAW_FP64)
    rp, nb, c = ctx.graph_csr(g)
    plain = ctx.solve(g)
    st, res = _compare(ctx, orc, rp, nb, c, "raw20k forced", exact=2)
    assert st["exact_certified"] == 1 and st["exact_gap"] <= 1e-10 * st["objective"]
    assert abs(st["objective"] - plain["objective"]) <= 1e-10 * plain["objective"]
    _feasible(rp, nb, res["x"])


def test_exact_cold_potentials_schedule(ctx, orc, golden):
    # stop the reference early (after the warm start): the exact stage then
    # starts far from the optimum and must still agree bit for bit
    case = [c for c in golden["lp"] if c["name"] == "C1"][0]
    x, y, mode, K = W.c1()
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    st, res = _compare(ctx, orc, rp, nb, c, "C1 glob_rounds=1", exact=2, glob_rounds=1, max_outer=1)
    assert st["objective"] == case["lp_opt"]
    _feasible(rp, nb, res["x"])


def _csr(n, edges):
    adj = [[] for _ in range(n)]
    for u, w, cst in edges:
        adj[u].append((cst, w))
        adj[w].append((cst, u))
    rp, nb, cc = [0], [], []
    for v in range(n):
        for cst, w in sorted(adj[v]):
            nb.append(w)
            cc.append(cst)
        rp.append(len(nb))
    return np.array(rp, np.int32), np.array(nb, np.int32), np.array(cc, np.float64)


def test_exact_error_paths(ctx, orc):
    # infeasible LP (bowtie): the round cap ends the stage, status 1.  The
    # reference schedule never converges on an infeasible LP (its dual is
    # unbounded), so it is cut to one round of each phase here.
    rp, nb, c = _csr(5, [(0, 1, 1.0), (0, 2, 2.0), (1, 2, 1.0), (0, 3, 1.0), (0, 4, 3.0), (3, 4, 1.0)])
    st, _ = _compare(ctx, orc, rp, nb, c, "bowtie", exact=2, exact_max_rounds=500, glob_rounds=1, max_outer=1,
                     repair=0)
    assert st["status"] == 1 and st["exact_certified"] == 0 and st["exact_rounds"] == 500
    # asymmetric costs: rejected, status 1
    rp, nb, c = _csr(6, [(0, 1, 1.0), (1, 2, 1.0), (0, 2, 1.0), (3, 4, 1.0), (4, 5, 1.0), (3, 5, 1.0)])
    c = c.copy()
    c[0] = 1.5
    from f2m_amd.solver import default_opts
    g = ctx.graph_from_csr(rp, nb, c)
    say("exact asymmetric")
    st = ctx.solve(g, default_opts(exact=2))
    r = orc.solve(rp, nb, c, orc.default_opts(exact=2))
    assert st["status"] == r["stats"]["status"] == 1 and st["exact_ran"] == r["stats"]["exact_ran"] == 1
    # malformed caller CSRs are refused before any kernel reads them
    for bad_nb in (np.array([1, 2, 0, 9, 0, 1], np.int32), np.array([1, 2, 0, 1, 0, 2], np.int32)):
        with pytest.raises(RuntimeError):
            ctx.graph_from_csr(np.array([0, 2, 4, 6], np.int32), bad_nb, np.ones(6))
    with pytest.raises(RuntimeError):
        ctx.graph_from_csr(np.array([0, 4, 2, 6], np.int32), np.array([1, 2, 0, 2, 0, 1], np.int32), np.ones(6))


def test_exact_c3_full_lattice(ctx, golden):
    """C3 (n = 1e6 lattice, K = 10, NINT): the reference recovery fails on its
    ties; the default solve returns the HiGHS optimum exactly."""
    import time
    e = [e for e in golden["lattice"] if e["n"] == 1_000_000][0]
    x, y, mode, K = W.c3()
    g = ctx.knn_build(x, y, K, mode)
    say("C3 solve")
    t = time.time()
    st = ctx.solve(g)
    dt = time.time() - t
    res = ctx.result()
    rp, nb, c = ctx.graph_csr(g)
    print(f"C3 solve {dt:.2f}s", {k: st[k] for k in EXACT_KEYS}, flush=True)
    assert st["status"] == 0 and st["exact_ran"] == 1 and st["exact_certified"] == 1
    assert st["objective"] == e["lp_opt"]
    _feasible(rp, nb, res["x"])
    P, D = certificate(rp, nb, c, res["x"], res["F"], st["d_min"])
    assert P == e["lp_opt"] and D <= e["lp_opt"] and P - D < 0.5
