"""GPU tests of the batched persistent solve and the differentiable bound.

Bar: every instance of a batch gives the same graph, potentials, components,
x and schedule counts as a plain f2m_solve of that instance, bit for bit,
the exact recovery included (objective within 1e-12 relative: the
per-instance sums are reduced in a different order); every lanes-per-node
schedule gives identical bits, with the potentials of the sweeps in global or
shared memory; graph replays equal the eager result.  Backward: within 1e-4 relative of the FP64
CPU gradient and of central finite differences (eps = 1e-4) on instances
checked to be nondegenerate (same kNN graph and same x at +-eps).
"""
import os

import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu


def mixed_instances(seed=5, count=24):
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(100), np.log(3000))))
        p = (rng.random((n, 2), dtype=np.float32) * np.float32(1e6)).astype(np.float32)
        x, y = W.dedup(p[:, 0], p[:, 1])
        xs.append(x)
        ys.append(y)
    return xs, ys


@pytest.fixture(scope="module")
def batch_case():
    return mixed_instances()


def _slices(node_off, rowptr, i):
    a, b = node_off[i], node_off[i + 1]
    return a, b, rowptr[a], rowptr[b]


@pytest.mark.parametrize("mode", [W.COST_NINT, W.COST_RAW_FP64])
def test_batch_matches_plain_solves(ctx, batch_case, mode, monkeypatch):
    from f2m_amd.solver import Context
    xs, ys = batch_case
    say("batch build", len(xs), mode)
    b = ctx.knn_build_batched(xs, ys, 16, mode)
    node_off, rowptr, nbr, cost = ctx.batch_csr(b)
    say("batch solve")
    sts = ctx.solve_batched(b)
    res = ctx.batch_result(b)
    plain = Context(0)
    for i in range(len(xs)):
        a, e, sa, se = _slices(node_off, rowptr, i)
        g = plain.knn_build(xs[i], ys[i], 16, mode)
        rp, nb, c = plain.graph_csr(g)
        assert np.array_equal(rowptr[a:e + 1] - sa, rp) and np.array_equal(nbr[sa:se], nb)
        assert np.array_equal(cost[sa:se], c)
        say("plain solve", i, len(xs[i]))
        st = plain.solve(g)
        r = plain.result()
        sb = sts[i]
        for key in ("sweeps", "glob_rounds", "outer_rounds", "n_combs", "n_irregular_a", "n_irregular_b",
                    "status", "max_rank", "n_repairs", "exact_ran", "exact_phases", "exact_rounds",
                    "exact_certified", "exact_gap"):
            assert sb[key] == st[key], (i, key, sb[key], st[key])
        assert np.array_equal(res["F"][a:e], r["F"]), i
        if st["n_irregular_a"] == 0 or st["exact_ran"]:
            assert np.array_equal(res["x"][sa:se], r["x"]), i
        else:
            assert np.allclose(res["x"][sa:se], r["x"], atol=1e-6), i
        assert abs(sb["objective"] - st["objective"]) <= 1e-12 * abs(st["objective"])
        assert abs(res["objective"][i] - st["objective"]) <= 1e-12 * abs(st["objective"])
    plain.close()


def same_stats(a, b):
    """Schedules and components identical; the FP64 sums (objective, dual) may
    differ in the last bits across workgroup sizes (reduction width)."""
    for sa, sb in zip(a, b):
        for k in sa:
            if k in ("objective", "dual"):
                assert abs(sa[k] - sb[k]) <= 1e-12 * abs(sb[k]), (k, sa[k], sb[k])
            else:
                assert sa[k] == sb[k], (k, sa[k], sb[k])
    assert len(a) == len(b)


def test_batch_schedules_identical_and_graph_replay(ctx, batch_case, monkeypatch):
    xs, ys = batch_case
    b = ctx.knn_build_batched(xs, ys, 16, W.COST_NINT)
    ref = None
    # potentials in shared memory (two buffers where they fit, one buffer
    # refreshed from global memory after each sweep) or in global memory;
    # active lists in shared memory where they fit after the potentials (3)
    for lpn, bs, lds in (("1", "256", "2"), ("2", "256", "0"), ("4", "512", "1"), ("8", "1024", "0"),
                         ("2", "1024", "3"), ("4", "256", "1"), ("4", "256", "3"), ("4", "1024", "2")):
        monkeypatch.setenv("F2M_BATCH_LPN", lpn)
        monkeypatch.setenv("F2M_BATCH_BS", bs)
        monkeypatch.setenv("F2M_BATCH_LDS", lds)
        say("batch lpn", lpn, "threads", bs, "shared potentials", lds)
        sts = ctx.solve_batched(b)
        res = ctx.batch_result(b)
        if ref is None:
            ref = (sts, res)
        else:
            same_stats(sts, ref[0])
            assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    monkeypatch.delenv("F2M_BATCH_LPN")
    monkeypatch.delenv("F2M_BATCH_BS")
    monkeypatch.delenv("F2M_BATCH_LDS")
    # runs of sweeps with the lists held in registers (the default where the
    # instance fits two nodes per thread) against the pass that reads them
    # from shared memory every sweep, with the owners' halves read from shared memory (1) or from
    # registers (2, the default)
    for regs in ("0", "1", "2"):
        monkeypatch.setenv("F2M_BATCH_REGS", regs)
        say("batch sweeps with lists in registers", regs)
        sts = ctx.solve_batched(b)
        res = ctx.batch_result(b)
        same_stats(sts, ref[0])
        assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    monkeypatch.delenv("F2M_BATCH_REGS")
    # partitions with a thread per node against four lanes per node (the default)
    monkeypatch.setenv("F2M_BATCH_PART_LANES", "0")
    say("batch partitions with a thread per node")
    sts = ctx.solve_batched(b)
    res = ctx.batch_result(b)
    same_stats(sts, ref[0])
    assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    monkeypatch.delenv("F2M_BATCH_PART_LANES")
    # captured graph: first call captures, later calls replay; eager work between
    outs = []
    for rep in range(3):
        say("graph replay", rep)
        sts = ctx.solve_batched(b)
        res = ctx.batch_result(b)
        outs.append((sts, res))
        _ = ctx.knn_directed(xs[0], ys[0], 5, W.COST_NINT)  # unrelated eager work on the stream
    for sts, res in outs:
        assert sts == outs[0][0]  # replays of one captured graph: identical bits
        same_stats(sts, ref[0])
        assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    # the size buckets of this batch launched on one stream, or on two
    for ov in ("0", "1"):
        monkeypatch.setenv("F2M_BATCH_OVERLAP", ov)
        say("buckets overlapped on", ov, "side streams")
        sts = ctx.solve_batched(b)
        res = ctx.batch_result(b)
        assert sts == outs[0][0], ov
        assert np.array_equal(res["F"], ref[1]["F"]) and np.array_equal(res["x"], ref[1]["x"])
    monkeypatch.delenv("F2M_BATCH_OVERLAP")
    monkeypatch.setenv("F2M_NO_GRAPH", "1")
    sts = ctx.solve_batched(b)
    assert sts == outs[0][0]


def test_batch_autotune_keeps_results(ctx, batch_case, tmp_path):
    xs, ys = batch_case
    b = ctx.knn_build_batched(xs, ys, 16, W.COST_NINT)
    ref = ctx.solve_batched(b)
    cache = str(tmp_path / "tune.txt")
    say("autotune")
    ctx.batch_autotune(b, sample=len(xs), cache_path=cache)
    lines = open(cache).read().split()
    assert len(lines) == 28 and set(lines[1::4]) <= {"1", "2", "4", "8"} and set(lines[2::4]) <= {"256", "512", "1024"}
    assert set(lines[3::4]) == {"-1"}
    same_stats(ctx.solve_batched(b), ref)
    ctx.batch_autotune(b, sample=4, cache_path=cache)  # loads the cache


def _nondegenerate_fd(orc, x, y, K, u, h):
    """Central FD of the F2M optimum (FP64 oracle, RAW_FP64 costs) at coordinate u,
    or None if the kNN graph or x changes within +-h."""
    vals = []
    ref = None
    for sgn in (+1, -1):
        xx = x.copy()
        xx[u] += sgn * h
        rp, nb, c = orc.graph_build(xx, y, K, W.COST_RAW_FP64)
        r = orc.solve(rp, nb, c)
        key = (rp.tobytes(), nb.tobytes(), r["x"].tobytes())
        if ref is None:
            ref = key
        elif key != ref or r["stats"]["n_irregular_a"]:
            return None
        vals.append(r["stats"]["objective"])
    return (vals[0] - vals[1]) / (2 * h)


def test_bound_backward_fp64_and_fd(orc):
    import torch

    from f2m_amd.autograd import F2MBound
    from f2m_amd.solver import Context
    solver = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    insts, _, K = W.c5(count=4, n=400)
    xs = [a for a, _ in insts]
    ys = [b for _, b in insts]
    node_off = np.zeros(len(xs) + 1, np.int64)
    node_off[1:] = np.cumsum([len(a) for a in xs])
    px = torch.tensor(np.concatenate(xs), dtype=torch.float32, device="cuda", requires_grad=True)
    py = torch.tensor(np.concatenate(ys), dtype=torch.float32, device="cuda", requires_grad=True)
    say("autograd forward")
    V = F2MBound.apply(px, py, node_off, K, solver, None)
    w = torch.arange(1, len(xs) + 1, dtype=torch.float64, device="cuda")
    say("autograd backward")
    (V * w).sum().backward()
    gx = px.grad.cpu().numpy()
    gy = py.grad.cpu().numpy()
    # FP64 CPU gradient with the GPU's x on the same graph
    node_off_b, rowptr, nbr, cost = solver.batch_csr()
    res = solver.batch_result()
    for i in range(len(xs)):
        a, e = node_off[i], node_off[i + 1]
        sa, se = rowptr[a], rowptr[e]
        rp = (rowptr[a:e + 1] - sa).astype(np.int32)
        ex, ey = orc.bound_backward(rp, nbr[sa:se], res["x"][sa:se], xs[i], ys[i])
        scale = float(i + 1)
        gref = np.concatenate([ex, ey]) * scale
        gget = np.concatenate([gx[a:e], gy[a:e]])
        assert np.linalg.norm(gget - gref) <= 1e-4 * np.linalg.norm(gref), i
        # objective agrees with an FP64 oracle solve of the same RAW_FP32 graph
        r = orc.solve(rp, nbr[sa:se], cost[sa:se])
        assert abs(r["stats"]["objective"] - V[i].item()) <= 1e-9 * abs(r["stats"]["objective"])
    # central finite differences on the first instance (FP64 costs), eps = 1e-4
    x0, y0 = xs[0].astype(np.float64), ys[0].astype(np.float64)
    checked = 0
    for u in range(0, len(x0), 37):
        fd = _nondegenerate_fd(orc, x0, y0, K, u, 1e-4)
        if fd is None:
            continue
        checked += 1
        assert abs(fd - gx[u]) <= 1e-4 * max(1.0, abs(fd)), (u, fd, gx[u])
    assert checked >= 3
    solver.close()


def test_batch_exact_on_lattices(ctx, golden):
    """Tie-heavy lattices: the reference recovery fails, every instance runs
    the in-workgroup exact stage; bit for bit equal to plain solves, and the
    n = 2000 lattice gives the HiGHS optimum exactly."""
    from f2m_amd.solver import Context
    sizes = (500, 2000, 1200, 3000, 800)
    xs, ys = [], []
    for n in sizes:
        x, y, mode, K = W.c3(n=n)
        xs.append(x)
        ys.append(y)
    say("batch lattices", sizes)
    b = ctx.knn_build_batched(xs, ys, K, mode)
    node_off, rowptr, nbr, cost = ctx.batch_csr(b)
    sts = ctx.solve_batched(b)
    res = ctx.batch_result(b)
    plain = Context(0)
    for i, n in enumerate(sizes):
        a, e, sa, se = _slices(node_off, rowptr, i)
        g = plain.knn_build(xs[i], ys[i], K, mode)
        say("plain lattice", n)
        st = plain.solve(g)
        r = plain.result()
        sb = sts[i]
        assert sb["exact_ran"] == 1 and sb["exact_certified"] == 1 and sb["status"] == 0, (n, sb)
        for key in ("sweeps", "n_combs", "exact_phases", "exact_rounds", "exact_gap", "max_rank"):
            assert sb[key] == st[key], (n, key, sb[key], st[key])
        assert np.array_equal(res["F"][a:e], r["F"]) and np.array_equal(res["x"][sa:se], r["x"]), n
        assert sb["objective"] == st["objective"]  # half-integral x on integer costs: exact sums
        if n == 2000:
            opt = [l for l in golden["lattice"] if l["n"] == 2000][0]["lp_opt"]
            assert sb["objective"] == opt
    plain.close()


def test_batch_warm_start_matches_plain_warm(ctx, batch_case):
    from f2m_amd.solver import Context
    xs, ys = batch_case
    b = ctx.knn_build_batched(xs, ys, 16, W.COST_NINT)
    node_off, rowptr, nbr, cost = ctx.batch_csr(b)
    say("batch cold")
    cold = ctx.solve_batched(b)
    F = ctx.batch_result(b)["F"]
    Fw = F + 0.01 * np.sin(np.arange(len(F)))  # a nearby starting point
    say("batch warm")
    warm = ctx.solve_batched(b, F_init=Fw)
    res = ctx.batch_result(b)
    plain = Context(0)
    for i in range(0, len(xs), 5):
        a, e, sa, se = _slices(node_off, rowptr, i)
        g = plain.knn_build(xs[i], ys[i], 16, W.COST_NINT)
        say("plain warm", i)
        st = plain.solve(g, F_init=Fw[a:e])
        r = plain.result()
        for k in ("sweeps", "glob_rounds", "outer_rounds", "n_combs", "status", "exact_ran", "exact_rounds"):
            assert warm[i][k] == st[k], (i, k)
        assert np.array_equal(res["F"][a:e], r["F"]) and np.array_equal(res["x"][sa:se], r["x"])
        assert abs(warm[i]["objective"] - cold[i]["objective"]) <= 1e-9 * abs(cold[i]["objective"])
    plain.close()
    # replays of the warm graph with new starting values (same context buffer)
    warm2 = ctx.solve_batched(b, F_init=F)
    assert all(w["status"] == 0 for w in warm2)


def test_batch_plans_warm_cold_alternating(ctx, batch_case):
    """The context keeps a warm and a cold plan of a batch: alternating
    solves replay both graphs and give the first results every time, also
    after a larger batch has grown the workspaces in between."""
    xs, ys = batch_case
    xa, ya = xs[:9], ys[:9]
    b = ctx.knn_build_batched(xa, ya, 16, W.COST_NINT)
    cold = ctx.solve_batched(b)
    rc = ctx.batch_result(b)
    Fw = rc["F"] + 0.01 * np.cos(np.arange(len(rc["F"])))
    warm = ctx.solve_batched(b, F_init=Fw)
    rw = ctx.batch_result(b)
    for rep in range(2):
        say("alternating cold / warm", rep)
        assert ctx.solve_batched(b) == cold
        r = ctx.batch_result(b)
        assert np.array_equal(r["F"], rc["F"]) and np.array_equal(r["x"], rc["x"])
        assert ctx.solve_batched(b, F_init=Fw) == warm
        r = ctx.batch_result(b)
        assert np.array_equal(r["F"], rw["F"]) and np.array_equal(r["x"], rw["x"])
    xb, yb = mixed_instances(seed=11, count=40)
    say("larger batch between")
    bb = ctx.knn_build_batched(xb, yb, 16, W.COST_NINT)
    assert all(st["status"] == 0 for st in ctx.solve_batched(bb))
    b = ctx.knn_build_batched(xa, ya, 16, W.COST_NINT)
    say("first batch again, warm then cold")
    assert ctx.solve_batched(b, F_init=Fw) == warm
    r = ctx.batch_result(b)
    assert np.array_equal(r["F"], rw["F"]) and np.array_equal(r["x"], rw["x"])
    assert ctx.solve_batched(b) == cold
    r = ctx.batch_result(b)
    assert np.array_equal(r["F"], rc["F"]) and np.array_equal(r["x"], rc["x"])


def test_autograd_warm_start(orc):
    import torch

    from f2m_amd.autograd import F2MBound, WarmStart
    from f2m_amd.solver import Context
    solver = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    insts, _, K = W.c5(count=8, n=400)
    xs = np.concatenate([a for a, _ in insts])
    ys = np.concatenate([b for _, b in insts])
    node_off = np.arange(0, 8 * 400 + 1, 400, dtype=np.int64)
    warm = WarmStart()
    rng = np.random.default_rng(3)
    for step in range(3):
        jx = (xs + step * 1e-4 * rng.standard_normal(len(xs))).astype(np.float32)
        jy = (ys + step * 1e-4 * rng.standard_normal(len(ys))).astype(np.float32)
        px = torch.tensor(jx, device="cuda", requires_grad=True)
        py = torch.tensor(jy, device="cuda", requires_grad=True)
        say("autograd warm step", step)
        Vw = F2MBound.apply(px, py, node_off, K, solver, None, warm)
        Vw.sum().backward()
        Vc = F2MBound.apply(px.detach(), py.detach(), node_off, K, solver, None)
        assert torch.allclose(Vw, Vc, rtol=1e-9, atol=0), step
        assert warm.F is not None and warm.F.numel() == len(xs)
    solver.close()
