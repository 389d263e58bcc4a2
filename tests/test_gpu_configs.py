"""GPU parity at the batched and differentiable workloads' own sizes.

Batched (4096 instances, n = floor(exp(U[ln 100, ln 20000])), fp32 points
x 1e6, NINT costs, K=16): 32 of the 4096 instances, spread over the whole
size range with the largest ones included, in Morton numbering as the
benchmark solves them, packed into one f2m_solve_batched call.  Every
instance is compared with the oracle (orc_solve, the CPU restatement of the
reference schedule graph_cuda.cu:47-155) on the same CSR: potentials F, x
and comb_root bit for bit, the schedule and component counts equal, the
objective within 1e-9 relative; the CSR itself equals the oracle's kNN
graph of the instance.

Differentiable (256 instances of n=2000, fp32 points in [0,1)^2, RAW_FP32
costs, K=20): the first 8 instances through F2MBound.  The bound equals the
oracle optimum of the same graph (1e-9 relative), the gradient equals the
FP64 CPU gradient (orc_bound_backward, 1e-4 relative) and central finite
differences of the FP64 oracle optimum (eps = 1e-4) at coordinates where the
kNN graph and x do not change within +-eps.
"""
import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

ORC_THREADS = 16
COUNT_KEYS = ("sweeps", "glob_rounds", "outer_rounds", "n_combs", "n_irregular_a", "n_irregular_b", "n_repairs",
              "status", "max_rank", "exact_ran", "exact_phases", "exact_rounds", "exact_certified")


def c4_selection(count=32):
    """Instance indices: 24 at evenly spaced size quantiles, the 2 largest,
    and 6 more evenly spaced among those with n >= 10200 (where one potential
    buffer no longer fits beside the other in shared memory)."""
    sizes = W.c4_sizes()
    order = np.argsort(sizes, kind="stable")
    pick = list(order[np.linspace(0, len(order) - 3, 24).astype(int)])
    pick += list(order[-2:])
    big = [i for i in order if sizes[i] >= 10200 and i not in pick]
    pick += [big[j] for j in np.linspace(0, len(big) - 1, count - len(pick)).astype(int)]
    assert len(set(pick)) == count
    return [int(i) for i in pick], sizes


def test_c4_instances_batched_vs_oracle(ctx, orc):
    idx, sizes = c4_selection()
    assert sum(sizes[i] >= 10200 for i in idx) >= 4 and sum(sizes[i] >= 19500 for i in idx) >= 2
    insts, mode, K = W.c4(only=idx)
    xs = [a for a, _ in insts]
    ys = [b for _, b in insts]
    perms = ctx.spatial_order_batched(xs, ys)
    xs = [a[p] for a, p in zip(xs, perms)]
    ys = [b[p] for b, p in zip(ys, perms)]
    say("c4 batch build", len(xs), "instances, n", min(map(len, xs)), "..", max(map(len, xs)))
    b = ctx.knn_build_batched(xs, ys, K, mode)
    node_off, rowptr, nbr, cost = ctx.batch_csr(b)
    say("c4 batch solve")
    sts = ctx.solve_batched(b)
    res = ctx.batch_result(b, comb_root=True)
    for i in range(len(xs)):
        a, e = node_off[i], node_off[i + 1]
        sa, se = rowptr[a], rowptr[e]
        rp = (rowptr[a:e + 1] - sa).astype(np.int32)
        ork = orc.graph_build(xs[i], ys[i], K, mode, threads=ORC_THREADS)
        assert np.array_equal(rp, ork[0]) and np.array_equal(nbr[sa:se], ork[1]) and np.array_equal(cost[sa:se], ork[2])
        say("c4 oracle", idx[i], "n", e - a)
        r = orc.solve(rp, nbr[sa:se], cost[sa:se], threads=ORC_THREADS)
        ro = r["stats"]
        for key in COUNT_KEYS:
            assert sts[i][key] == ro[key], (idx[i], key, sts[i][key], ro[key])
        assert np.array_equal(res["F"][a:e], r["F"]), idx[i]
        assert np.array_equal(res["x"][sa:se], r["x"]), idx[i]
        assert np.array_equal(res["comb_root"][a:e], r["comb_root"]), idx[i]
        assert sts[i]["status"] == 0
        assert abs(sts[i]["objective"] - ro["objective"]) <= 1e-9 * abs(ro["objective"])
        assert abs(res["objective"][i] - ro["objective"]) <= 1e-9 * abs(ro["objective"])


def _fd(orc, x, y, K, u, h, axis):
    """Central FD of the FP64 oracle optimum (RAW_FP64 costs) in coordinate u
    of `axis` (0: x, 1: y), or None where the kNN graph or x changes within
    +-h or the tight components are irregular (x* not unique)."""
    vals = []
    ref = None
    for sgn in (+1, -1):
        xx, yy = x.copy(), y.copy()
        (xx if axis == 0 else yy)[u] += sgn * h
        rp, nb, c = orc.graph_build(xx, yy, K, W.COST_RAW_FP64, threads=ORC_THREADS)
        r = orc.solve(rp, nb, c, threads=ORC_THREADS)
        key = (rp.tobytes(), nb.tobytes(), r["x"].tobytes())
        if ref is None:
            ref = key
        elif key != ref or r["stats"]["n_irregular_a"]:
            return None
        vals.append(r["stats"]["objective"])
    return (vals[0] - vals[1]) / (2 * h)


def test_c5_instances_bound_and_gradient_vs_oracle(orc):
    import torch

    from f2m_amd.autograd import F2MBound
    from f2m_amd.solver import Context
    solver = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    insts, mode, K = W.c5(count=8)  # the first 8 of the 256 instances
    xs = [a for a, _ in insts]
    ys = [b for _, b in insts]
    assert all(len(a) >= 1990 for a in xs)
    node_off = np.zeros(len(xs) + 1, np.int64)
    node_off[1:] = np.cumsum([len(a) for a in xs])
    px = torch.tensor(np.concatenate(xs), dtype=torch.float32, device="cuda", requires_grad=True)
    py = torch.tensor(np.concatenate(ys), dtype=torch.float32, device="cuda", requires_grad=True)
    say("c5 forward 8 x n=2000")
    V = F2MBound.apply(px, py, node_off, K, solver, None)
    w = torch.linspace(0.5, 2.0, len(xs), dtype=torch.float64, device="cuda")
    say("c5 backward")
    (V * w).sum().backward()
    gx = px.grad.cpu().numpy()
    gy = py.grad.cpu().numpy()
    Vh = V.detach().cpu().numpy()
    wh = w.cpu().numpy()
    _, rowptr, nbr, cost = solver.batch_csr()
    res = solver.batch_result()
    for i in range(len(xs)):
        a, e = node_off[i], node_off[i + 1]
        sa, se = rowptr[a], rowptr[e]
        rp = (rowptr[a:e + 1] - sa).astype(np.int32)
        ork = orc.graph_build(xs[i], ys[i], K, mode, threads=ORC_THREADS)
        assert np.array_equal(rp, ork[0]) and np.array_equal(nbr[sa:se], ork[1]) and np.array_equal(cost[sa:se], ork[2])
        say("c5 oracle", i)
        r = orc.solve(rp, nbr[sa:se], cost[sa:se], threads=ORC_THREADS)
        assert r["stats"]["status"] == 0
        assert abs(r["stats"]["objective"] - Vh[i]) <= 1e-9 * abs(r["stats"]["objective"]), i
        assert np.array_equal(res["x"][sa:se], r["x"]), i
        ex, ey = orc.bound_backward(rp, nbr[sa:se], r["x"], xs[i], ys[i])
        gref = np.concatenate([ex, ey]) * wh[i]
        gget = np.concatenate([gx[a:e], gy[a:e]])
        assert np.linalg.norm(gget - gref) <= 1e-4 * np.linalg.norm(gref), i
    # central finite differences (eps = 1e-4) on instance 0, both axes
    x0, y0 = xs[0], ys[0]
    checked = 0
    for u in range(0, len(x0), 97):
        for axis in (0, 1):
            fd = _fd(orc, x0, y0, K, u, 1e-4, axis)
            if fd is None:
                continue
            checked += 1
            got = (gx if axis == 0 else gy)[u] / wh[0]
            assert abs(fd - got) <= 1e-4 * max(1.0, abs(fd)), (u, axis, fd, got)
    say("c5 finite differences checked", checked)
    assert checked >= 10
    solver.close()


def test_autograd_pending_forwards_keep_their_graphs(orc):
    """Forward A, forward B (larger, same context), backward A, backward B:
    each gradient is that of its own graph and x."""
    import torch

    from f2m_amd.autograd import F2MBound
    from f2m_amd.solver import Context
    solver = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    cases = []
    for count, n, seed in ((3, 300, 41), (5, 900, 42)):
        insts, mode, K = W.c5(count=count, n=n, seed=seed)
        xs = [a for a, _ in insts]
        ys = [b for _, b in insts]
        off = np.zeros(count + 1, np.int64)
        off[1:] = np.cumsum([len(a) for a in xs])
        px = torch.tensor(np.concatenate(xs), dtype=torch.float32, device="cuda", requires_grad=True)
        py = torch.tensor(np.concatenate(ys), dtype=torch.float32, device="cuda", requires_grad=True)
        say("pending forward", count, n)
        V = F2MBound.apply(px, py, off, K, solver, None)
        cases.append((xs, ys, off, px, py, V, K, mode))
    for xs, ys, off, px, py, V, K, mode in cases:  # A's backward after B's forward
        say("pending backward", len(xs))
        V.sum().backward()
        gx = px.grad.cpu().numpy()
        gy = py.grad.cpu().numpy()
        for i in range(len(xs)):
            rp, nb, c = orc.graph_build(xs[i], ys[i], K, mode, threads=ORC_THREADS)
            r = orc.solve(rp, nb, c, threads=ORC_THREADS)
            assert abs(r["stats"]["objective"] - V[i].item()) <= 1e-9 * abs(r["stats"]["objective"])
            ex, ey = orc.bound_backward(rp, nb, r["x"], xs[i], ys[i])
            gref = np.concatenate([ex, ey])
            gget = np.concatenate([gx[off[i]:off[i + 1]], gy[off[i]:off[i + 1]]])
            assert np.linalg.norm(gget - gref) <= 1e-4 * np.linalg.norm(gref), (len(xs), i)
    solver.close()
