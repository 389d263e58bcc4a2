"""GPU parity tests: the HIP path (through the C ABI) against the oracle.

Bar: bit-exact for kNN ids, graph CSR (ids and FP64 costs), every sweep's F,
the potentials after each schedule phase, tight-component roots and x where
the tight components are all regular; the objective within 1e-9 relative of
the oracle and of the HiGHS LP optimum (tests/golden/golden.json).
"""
import os

import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.conftest import say
from tests.helpers import feasibility, graph_hash, instance_from_spec, sym_check

pytestmark = pytest.mark.gpu


def test_context_and_version(ctx):
    from f2m_amd import _lib as L
    say("ctx")
    ctx.synchronize()
    assert L.lib.f2m_version().startswith(b"f2m_hip")


@pytest.mark.parametrize("name,n,span,K,mode,seed", [
    ("c1", None, None, 20, W.COST_NINT, None),
    ("ties", 3000, 60, 10, W.COST_NINT, 11),
    ("raw64", 20000, 10**7, 20, W.COST_RAW_FP64, 12),
    ("fp32", 5000, 10**6, 16, W.COST_RAW_FP32, 13),
    ("k50", 4000, 10**5, 50, W.COST_RAW_FP64, 14),
])
def test_knn_directed_matches_oracle(ctx, orc, name, n, span, K, mode, seed):
    if name == "c1":
        x, y, mode, K = W.c1()
    else:
        x, y = W.uniform_int(n, span, seed)
        if mode == W.COST_RAW_FP32:
            x = (x / span).astype(np.float32).astype(np.float64)
            y = (y / span).astype(np.float32).astype(np.float64)
        if mode == W.COST_RAW_FP64:
            x = x + np.random.default_rng(seed).random(len(x))  # non-integer coordinates
    say("knn_directed", name, len(x), K, mode)
    a = ctx.knn_directed(x, y, K, mode)
    b = orc.knn_directed(x, y, K, mode)
    assert np.array_equal(a, b)


def test_graph_matches_fixtures_and_oracle(ctx, orc, golden):
    for case in golden["lp"]:
        x, y = instance_from_spec(case["spec"])
        say("knn_build", case["name"])
        g = ctx.knn_build(x, y, case["K"], case["mode"])
        rp, nb, c = ctx.graph_csr(g)
        assert graph_hash(rp, nb, c) == case["graph_sha256"], case["name"]


def test_graph_c2_matches_oracle(ctx, orc):
    x, y, mode, K = W.c2()
    say("knn_build C2")
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    r2 = orc.graph_build(x, y, K, mode)
    assert np.array_equal(rp, r2[0]) and np.array_equal(nb, r2[1]) and np.array_equal(c, r2[2])


@pytest.mark.parametrize("lpn", [1, 2, 4, 8, 16, 64])
def test_sweep_bit_exact(ctx, orc, lpn, monkeypatch):
    x, y, mode, K = W.c1()
    rp, nb, c = orc.graph_build(x, y, K, mode)
    D = c / c[c > 0].min()
    rng = np.random.default_rng(lpn)
    n = len(rp) - 1
    deg = np.diff(rp)
    length = np.minimum(deg, rng.integers(0, 40, size=n)).astype(np.int32)
    F = rng.normal(scale=3.0, size=n)
    monkeypatch.setenv("F2M_LPN", str(lpn))
    say("sweep lpn", lpn)
    Fg, gg, mx = ctx.sweep(rp, nb, D, length, F, nsweeps=1)
    Fo, go = orc.sweep(rp, nb, D, length, F)
    assert np.array_equal(Fg, Fo) and np.array_equal(gg, go)
    assert mx == np.abs(go).max()
    # several sweeps in a row
    Fg5, _, _ = ctx.sweep(rp, nb, D, length, F, nsweeps=5)
    Fo5 = F
    for _ in range(5):
        Fo5, _ = orc.sweep(rp, nb, D, length, Fo5)
    assert np.array_equal(Fg5, Fo5)


@pytest.mark.parametrize("phase", [1, 2])
def test_schedule_phases_bit_exact(ctx, orc, phase):
    from f2m_amd.solver import default_opts
    x, y, mode, K = W.c1()
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    say("solve stop_phase", phase)
    st = ctx.solve(g, default_opts(stop_phase=phase))
    res = ctx.result()
    r = orc.solve(rp, nb, c, orc.default_opts(stop_phase=phase))
    assert st["sweeps"] == r["stats"]["sweeps"]
    assert st["glob_rounds"] == r["stats"]["glob_rounds"]
    assert np.array_equal(res["F"], r["F"])


@pytest.mark.parametrize("rows", ["0", "1", "2"])
def test_single_sweeps_after_tile_builds(ctx, orc, monkeypatch, rows):
    # one sweep per block_GLOB round and block_EF chunks of 7 (remainders of one
    # sweep): single sweeps run the padded kernel on lists whose tiles were
    # just built, so the padded lists are built for them on demand when the
    # tile build reads the rows (F2M_BUILD_ROWS 1, 2)
    from f2m_amd.solver import default_opts
    monkeypatch.setenv("F2M_BUILD_ROWS", rows)
    x, y = W.uniform_int(5000, 10**6, 29)
    g = ctx.knn_build(x, y, 12, W.COST_NINT)
    rp, nb, c = ctx.graph_csr(g)
    kw = dict(glob_sweeps=1, glob_rounds=40, check_every=7)
    say("solve uniform5000 single sweeps, F2M_BUILD_ROWS", rows)
    st = ctx.solve(g, default_opts(**kw))
    res = ctx.result()
    r = orc.solve(rp, nb, c, orc.default_opts(**kw))
    so = r["stats"]
    for k in ("sweeps", "glob_rounds", "outer_rounds", "max_grad", "status"):
        assert st[k] == so[k], (k, st[k], so[k])
    assert np.array_equal(res["F"], r["F"]) and np.array_equal(res["x"], r["x"])


def _full_compare(ctx, orc, x, y, K, mode, name, lp_opt=None, exact_x=True):
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    say("solve", name, len(x), K, mode)
    st = ctx.solve(g)
    res = ctx.result()
    r = orc.solve(rp, nb, c)
    so = r["stats"]
    assert st["sweeps"] == so["sweeps"], (st, so)
    assert np.array_equal(res["F"], r["F"]), name
    assert np.array_equal(res["comb_root"], r["comb_root"]), name
    assert st["n_combs"] == so["n_combs"] and st["n_irregular_a"] == so["n_irregular_a"]
    assert st["status"] == so["status"] == 0
    assert abs(st["objective"] - so["objective"]) <= 1e-9 * abs(so["objective"])
    assert abs(st["dual"] - so["dual"]) <= 1e-9 * abs(so["dual"])
    assert st["max_rank"] == so["max_rank"]
    for k in ("exact_ran", "exact_phases", "exact_rounds", "exact_certified", "exact_gap"):
        assert st[k] == so[k], (name, k, st[k], so[k])
    if lp_opt is not None:
        assert abs(st["objective"] - lp_opt) <= 1e-9 * abs(lp_opt)
    if so["exact_ran"] or (exact_x and so["n_irregular_a"] == 0):
        assert np.array_equal(res["x"], r["x"])
    else:
        # irregular components: least squares on the GPU sums in another order
        assert np.allclose(res["x"], r["x"], atol=1e-6)
    assert sym_check(rp, nb, res["x"])
    return st, res, rp, nb


def test_full_solve_golden_cases(ctx, orc, golden):
    for case in golden["lp"]:
        x, y = instance_from_spec(case["spec"])
        _full_compare(ctx, orc, x, y, case["K"], case["mode"], case["name"], lp_opt=case["lp_opt"])


def test_full_solve_c2(ctx, orc):
    x, y, mode, K = W.c2()
    st, res, rp, nb = _full_compare(ctx, orc, x, y, K, mode, "C2")
    err, lo, hi = feasibility(rp, nb, res["x"])
    assert err < 1e-6


def test_full_solve_c2_morton_every_sweep_width(ctx, orc, monkeypatch):
    """Morton-numbered C2: the schedule's sweeps run several per launch on
    spatial tiles (halo levels recomputed in shared memory).  The default
    width equals the oracle bit for bit, and every width 1..12 gives the same
    potentials and sweep count; so do other tile sizes, workgroup mappings
    and tile boundaries."""
    x0, y0, mode, K = W.c2()
    perm = ctx.spatial_order(x0, y0)
    x, y = x0[perm], y0[perm]
    st, res, rp, nb = _full_compare(ctx, orc, x, y, K, mode, "C2 Morton")
    p = ctx.profile()
    assert p["sweeps"] == st["sweeps"] and p["sweep_launches"] < st["sweeps"] / 3, p
    g = ctx.knn_build(x, y, K, mode)
    for w in (1, 2, 3, 4, 5, 6, 8, 12):
        monkeypatch.setenv("F2M_FUSE", str(w))
        say("solve C2 Morton, F2M_FUSE", w)
        s2 = ctx.solve(g)
        p = ctx.profile()
        assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"], w
        assert np.array_equal(ctx.result()["F"], res["F"]), w
        assert (p["sweep_launches"] == st["sweeps"]) if w == 1 else (p["sweep_launches"] < st["sweeps"]), (w, p)
    monkeypatch.delenv("F2M_FUSE")
    # block_GLOB's partition with one thread per node instead of four lanes
    monkeypatch.setenv("F2M_PART_LANES", "1")
    say("solve C2 Morton, F2M_PART_LANES 1")
    s2 = ctx.solve(g)
    assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"]
    assert np.array_equal(ctx.result()["F"], res["F"]) and np.array_equal(ctx.result()["x"], res["x"])
    monkeypatch.delenv("F2M_PART_LANES")
    # block_GLOB tile sweeps reading the owner's half from shared memory instead of a register;
    # the convergence checks' gradient from a separate pass instead of the tile build; the tile
    # build reading the full padded lists (ROWS 0) or the rows only (ROWS 2); every build probing
    # the halo level past the last fit (PROBE 1)
    for var, val in (("F2M_TILE_SELH", "0"), ("F2M_BUILD_GRAD", "0"), ("F2M_BUILD_ROWS", "0"),
                     ("F2M_BUILD_ROWS", "2"), ("F2M_TILE_PROBE", "1")):
        monkeypatch.setenv(var, val)
        say("solve C2 Morton,", var, val)
        s2 = ctx.solve(g)
        assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"], var
        assert s2["glob_rounds"] == st["glob_rounds"] and s2["outer_rounds"] == st["outer_rounds"], var
        assert s2["max_grad"] == st["max_grad"], var
        assert np.array_equal(ctx.result()["F"], res["F"]), var
        monkeypatch.delenv(var)
    # nodes per tile of the block_GLOB and block_EF lists; their workgroup shapes (384 threads
    # for 768 local nodes, 768 for 1536)
    for tg, te, wg, we in (("256", "256", "768", "384"), ("64", "128", "384", "768"), ("196", "196", "384", "384"),
                           ("300", "450", "768", "768")):
        monkeypatch.setenv("F2M_TILE_GLOB", tg)
        monkeypatch.setenv("F2M_TILE", te)
        monkeypatch.setenv("F2M_TILE_W_GLOB", wg)
        monkeypatch.setenv("F2M_TILE_W", we)
        say("solve C2 Morton, tiles", tg, te, wg, we)
        s2 = ctx.solve(g)
        assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"], (tg, te)
        assert np.array_equal(ctx.result()["F"], res["F"]), (tg, te)
    for v in ("F2M_TILE_GLOB", "F2M_TILE", "F2M_TILE_W_GLOB", "F2M_TILE_W"):
        monkeypatch.delenv(v)
    # tile -> workgroup mapping: tile = workgroup, or grouped per XCD (3 does not divide the tiles)
    for xcd in ("1", "3"):
        monkeypatch.setenv("F2M_TILE_XCD", xcd)
        say("solve C2 Morton, tiles grouped over", xcd, "XCDs")
        s2 = ctx.solve(g)
        assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"], xcd
        assert np.array_equal(ctx.result()["F"], res["F"]), xcd
    monkeypatch.delenv("F2M_TILE_XCD")
    # tile boundaries: uniform, balanced by work for both list kinds, and with 100-node tiles
    # balanced for both (the boundaries move every rebuild)
    for bal, te in (("0", None), ("2", None), ("2", "100")):
        monkeypatch.setenv("F2M_TILE_BALANCE", bal)
        if te:
            monkeypatch.setenv("F2M_TILE", te)
            monkeypatch.setenv("F2M_TILE_GLOB", te)
        say("solve C2 Morton, tile balance", bal, te)
        s2 = ctx.solve(g)
        assert s2["sweeps"] == st["sweeps"] and s2["objective"] == st["objective"], bal
        assert np.array_equal(ctx.result()["F"], res["F"]), bal


def test_full_solve_random_raw(ctx, orc):
    for seed in (21, 22):
        x, y = W.uniform_int(20000, 10**7, seed)
        st, res, rp, nb = _full_compare(ctx, orc, x, y, 10, W.COST_RAW_FP64, f"raw20k-{seed}")
        assert st["n_irregular_a"] == 0
        assert set(np.unique(res["x"]).tolist()) <= {0.0, 0.5, 1.0}
        err, lo, hi = feasibility(rp, nb, res["x"])
        assert err == 0.0


def test_solve_is_deterministic(ctx):
    x, y, mode, K = W.c1()
    g = ctx.knn_build(x, y, K, mode)
    say("determinism")
    a = ctx.solve(g)
    ra = ctx.result()
    b = ctx.solve(g)
    rb = ctx.result()
    assert a == b
    for k in ra:
        assert np.array_equal(ra[k], rb[k])


def test_spatial_order_matches_restatement(ctx):
    from tests.helpers import morton_order
    x, y, _, _ = W.c2()
    say("spatial order C2")
    assert np.array_equal(ctx.spatial_order(x, y), morton_order(x, y))
    # collinear points (zero extent on one axis), repeated quantised keys, tiny sets
    xl = np.arange(5000, dtype=float) * 0.5
    yl = np.zeros(5000)
    assert np.array_equal(ctx.spatial_order(xl, yl), morton_order(xl, yl))
    xs, ys = W.uniform_int(3000, 10**9, 31)
    assert np.array_equal(ctx.spatial_order(xs, ys), morton_order(xs, ys))
    assert np.array_equal(ctx.spatial_order(np.array([3.0]), np.array([1.0])), [0])
    # batched: each instance permuted within itself
    insts = [W.uniform_int(n, 10**6, 40 + i) for i, n in enumerate((100, 2500, 7, 900))]
    perms = ctx.spatial_order_batched([a for a, _ in insts], [b for _, b in insts])
    for (a, b), p in zip(insts, perms):
        assert np.array_equal(p, morton_order(a, b))
    # 600 instances (six radix passes over instance and Morton bits), many
    # repeated keys on a coarse grid
    rng = np.random.default_rng(5)
    insts = [W.uniform_int(int(n), 300, 500 + i) for i, n in enumerate(rng.integers(1, 700, size=600))]
    say("spatial order, 600 instances")
    perms = ctx.spatial_order_batched([a for a, _ in insts], [b for _, b in insts])
    for (a, b), p in zip(insts, perms):
        assert np.array_equal(p, morton_order(a, b))


def test_morton_numbering_same_optimum(ctx, orc, golden):
    # C1 renumbered along the curve: the same LP optimum, bit-exact with the
    # oracle on the renumbered points
    from tests.helpers import morton_order
    case = [c for c in golden["lp"] if c["name"] == "C1"][0]
    x, y, mode, K = W.c1()
    p = morton_order(x, y)
    _full_compare(ctx, orc, x[p], y[p], K, mode, "C1 morton", lp_opt=case["lp_opt"])


def test_warm_start_and_knn_certificate(ctx, orc, golden):
    from f2m_amd.solver import default_opts
    x, y, mode, K = W.c1()
    opt = [c for c in golden["lp"] if c["name"] == "C1"][0]["lp_opt"]
    g4 = ctx.knn_build(x, y, 4, mode)
    say("solve K=4")
    ctx.solve(g4)
    F4 = ctx.result()["F"]
    v4 = ctx.knn_certify(g4)
    rp4, nb4, c4 = ctx.graph_csr(g4)
    r4 = orc.solve(rp4, nb4, c4)
    assert np.array_equal(F4, r4["F"])
    assert v4 == orc.knn_certify(rp4, c4, 4, r4["stats"]["d_min"], r4["F"]) > 0
    g = ctx.knn_build(x, y, K, mode)
    rp, nb, c = ctx.graph_csr(g)
    say("warm solve K=20 from K=4")
    st = ctx.solve(g, F_init=F4)
    res = ctx.result()
    r = orc.solve(rp, nb, c, F_init=F4)
    for k in ("sweeps", "glob_rounds", "outer_rounds", "n_combs", "status", "exact_ran", "exact_rounds"):
        assert st[k] == r["stats"][k], k
    assert np.array_equal(res["F"], r["F"]) and np.array_equal(res["x"], r["x"])
    assert st["objective"] == opt
    assert ctx.knn_certify(g) == orc.knn_certify(rp, c, K, r["stats"]["d_min"], r["F"]) == 0
    # the warm entry point accepts device memory too
    import torch
    Fd = torch.tensor(F4, device="cuda")
    st2 = ctx.solve(g, F_init=Fd)
    assert st2["sweeps"] == st["sweeps"] and st2["objective"] == st["objective"]
    with pytest.raises(RuntimeError):  # a caller CSR has no k
        gc = ctx.graph_from_csr(rp, nb, c)
        ctx.solve(gc)
        ctx.knn_certify(gc)


def test_solve_adaptive_certifies_c1(ctx, golden):
    x, y, mode, K = W.c1()
    opt = [c for c in golden["lp"] if c["name"] == "C1"][0]["lp_opt"]
    say("adaptive K from 4")
    k, st, res, hist = ctx.solve_adaptive(x, y, 4, mode)
    ks = [h[0] for h in hist]
    assert ks[0] == 4 and ks == sorted(ks) and hist[-1][2] == 0
    assert all(h[2] > 0 for h in hist[:-1])
    # certified: the optimum over all pairs of points, equal to the K=20 graph's
    assert st["objective"] == opt and st["status"] == 0
    assert all(h[3] > 0 for h in hist)
