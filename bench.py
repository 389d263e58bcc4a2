#!/usr/bin/env python3
"""F2M solve benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one plain F2M solve (f2m_solve: dual descent with the reference
schedule, tight components, primal recovery, FP64 objective) of the
n=100,000 / K=20 clustered workload: 50-component Gaussian mixture, NINT
costs, seed 0xF2A2.  Nodes are numbered along the Morton curve of the points
(f2m_spatial_order, as the CLI does; the reference numbers nodes in hash
order), and the kNN graph is built once before the timed region, as the
reference's "Time:" excludes graph construction (graph_cuda.cu:51,131).  The
same solve in input order is timed as well (input_order).
Prints one JSON line.  `value` = ms per solve (lower is better).

The timed solves run without instrumentation.

roofline: the dominant kernel is the multi-sweep launch over spatial tiles
(s_sweep_tile, F2M_FUSE sweeps per launch, default 10, one 768-thread tile per CU).  Its algorithmic
bytes per launch are, per sweep it runs, 8 B per active slot (nbr int32 +
a 4-byte cost) + 20 B per node (list offset, F in, F out), counted on the device
for the timed launches; the average kernel duration comes from the HIP
events the library records at the start and end of every 16th such launch
(hipExtLaunchKernelGGL, on its stream) in two further, instrumented solves.
traffic is read from profiles/traffic.json (rocprofv3 FETCH_SIZE +
WRITE_SIZE per launch of that kernel, scaled by a copy kernel of known size)
when it was measured for the same kernel.

cpu_baseline: the oracle (the CPU restatement of the reference schedule,
oracle/f2m_oracle.cpp) solving the same graph with the host's threads, and
the same solve with nodes in input order.

batched: the 4096-instance workload (n log-uniform in [100, 20000], fp32
points x 1e6, NINT costs, K=16) solved by one replay of the captured
persistent launch (f2m_solve_batched), against a loop of plain GPU solves
timed on 16 instances spread over the size range (per-instance throughput,
identical optima checked on those 16).

fwd_bwd: the differentiable bound on 256 instances of n=2000 (fp32 points
in the unit square, RAW_FP32 costs, K=20): forward (kNN + batched solve) and
backward (f2m_bound_backward through torch.autograd); then, after one untimed
training-loop step, the next step (points moved by ~1e-4) forwarded warm
(WarmStart) and cold.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "F2M solve ms @n=100k,k=20 (plain vs batched); fwd+bwd ms; % HBM peak"
# schedule table of the 4096-instance batch, written by `bench.py --tune-batch`
TUNED_C4 = os.path.join(ROOT, "f2m_amd", "tuned", "batch_n100-20000_k16.txt")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sample-every", type=int, default=16)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the solves, copy a buffer of this many MiB (counter calibration)")
    ap.add_argument("--no-fwd-bwd", action="store_true")
    ap.add_argument("--no-input-order", action="store_true")
    ap.add_argument("--batch-order", choices=["morton", "input"], default="morton")
    ap.add_argument("--only", choices=["batched", "fwd_bwd"], default=None,
                    help="run only this part and print its JSON (for profiling runs)")
    ap.add_argument("--batch-table", default=TUNED_C4,
                    help="schedule table of the batched leg (f2m_batch_autotune cache file; 'none': built-in)")
    ap.add_argument("--tune-batch", default=None, metavar="PATH",
                    help="choose the batched leg's schedule table (tune_batch_table) and write it to PATH")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the F2M hot path shards by instance or by spatial region, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    from f2m_amd import workloads as W
    from f2m_amd.solver import Context

    if args.calibrate_copy_mib:
        import torch
        a = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy", args.calibrate_copy_mib, "MiB")
        del a, b

    if args.only:
        if args.only == "batched":
            c = Context(0)
            print(json.dumps({"batched": bench_batched(c, args.batch_order, args.batch_table, args.tune_batch)}),
                  flush=True)
            c.close()
        else:
            print(json.dumps({"fwd_bwd": bench_fwd_bwd()}), flush=True)
        return

    x0, y0, mode, K = W.c2()
    n = len(x0)
    ctx = Context(0)
    input_order = time_input_order(ctx, x0, y0, K, mode, args) if not args.no_input_order else None
    log("spatial order + kNN graph build", n, K)
    t = time.perf_counter()
    perm = ctx.spatial_order(x0, y0)
    x, y = x0[perm], y0[perm]
    order_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    g = ctx.knn_build(x, y, K, mode)
    ctx.synchronize()
    knn_ms = (time.perf_counter() - t) * 1e3
    for i in range(args.warmup):
        st = ctx.solve(g)
        log("warmup", i, "sweeps", st["sweeps"], "objective", st["objective"])
    # timed solves: no instrumentation (sampling events cost ~10 us per timed launch)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.solve(g)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    log("timed", args.steps, "solves:", ms, "ms/solve")

    # roofline of the dominant sweep kernel, from separate instrumented solves:
    # each run of back-to-back launches of the full width timed with one pair
    # of HIP events on the solver's stream (the gaps between them included),
    # their algorithmic bytes counted on the device
    ctx.set_profile(args.sample_every)
    prof_tot = {"sweep_launches": 0, "sweep_samples": 0, "sweep_ms_sampled": 0.0, "sweeps": 0,
                "sampled_sweeps": 0, "sampled_alg_bytes": 0.0}
    for i in range(2):
        ctx.solve(g)
        p = ctx.profile()
        for k in prof_tot:
            prof_tot[k] += p[k]
    ctx.set_profile(0)
    samples = max(1, prof_tot["sweep_samples"])
    launch_ms = prof_tot["sweep_ms_sampled"] / samples
    bytes_per_launch = prof_tot["sampled_alg_bytes"] / samples
    sweeps_per_launch = prof_tot["sampled_sweeps"] / samples
    # no full-width launch sampled (the halos of a tile configuration did not
    # fit the configured width): no rate to report
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else None
    # the sampled launches are those of the configured width (F2M_FUSE)
    fuse = int(round(sweeps_per_launch)) if launch_ms > 0 else int(os.environ.get("F2M_FUSE", "10"))
    tw = 384 if os.environ.get("F2M_TILE_W") == "384" else 768  # the block_EF tiles' workgroup shape
    kernel = f"s_sweep_tile<{fuse},false,4,{tw}>" if fuse > 1 else "s_sweep_pad<4,2,true,1,false,true>"
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        if tj.get("kernel") == kernel:
            traffic = tj.get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None, "traffic": traffic,
                "kernel": kernel, "avg_launch_us": round(launch_ms * 1e3, 3),
                "sweeps_per_launch": round(sweeps_per_launch, 2),
                "alg_bytes_per_launch": round(bytes_per_launch, 1),
                "launches_per_solve": prof_tot["sweep_launches"] / 2, "sweeps_per_solve": prof_tot["sweeps"] / 2}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        host = host_cpu()
        # the job's CPU share: the cores this process may run on, at most 16
        # (the share of one GPU on a shared host); then one core
        share = min(host["affinity_cpus"], 16)
        rp, nb, c = ctx.graph_csr(g)
        runs = {}
        for threads in sorted({share, 1}, reverse=True):
            log("cpu baseline: oracle solve with", threads, "threads")
            t = time.perf_counter()
            r = O.solve(rp, nb, c, threads=threads)
            runs[threads] = ((time.perf_counter() - t) * 1e3, r)
        best = min(runs, key=lambda k: runs[k][0])
        cpu_ms, r = runs[best]
        ok = all(abs(rr["stats"]["objective"] - st["objective"]) <= 1e-9 * abs(st["objective"])
                 for _, rr in runs.values())
        cpu = {"value": round(cpu_ms, 1), "unit": "ms", "cores": best, "kind": "port",
               "sample": f"one full solve of the same n={n} K={K} graph, nodes in Morton order "
                         f"({r['stats']['sweeps']} sweeps), objective agrees: {ok}",
               "speedup_gpu_vs_cpu": round(cpu_ms / ms, 1),
               "ms_by_threads": {str(k): round(v[0], 1) for k, v in runs.items()},
               "speedup_vs_1_thread": round(runs[1][0] / ms, 1), **host}
        if input_order is not None:
            # the same solve with nodes in input order (the reference numbers
            # nodes arbitrarily): the CPU loses its cache locality too
            rp0, nb0, c0 = O.graph_build(x0, y0, K, mode, threads=share)
            t = time.perf_counter()
            r0 = O.solve(rp0, nb0, c0, threads=share)
            cpu0 = (time.perf_counter() - t) * 1e3
            cpu["input_order_ms"] = round(cpu0, 1)
            cpu["input_order_speedup"] = round(cpu0 / input_order["ms"], 1)
            cpu["input_order_objective_agrees"] = bool(
                abs(r0["stats"]["objective"] - input_order["objective"]) <= 1e-9 * abs(input_order["objective"]))

    batched = None if args.no_batched else bench_batched(ctx, args.batch_order, args.batch_table, args.tune_batch)
    fwd_bwd = None if args.no_fwd_bwd else bench_fwd_bwd()

    out = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain F2M solve, n=100000 50-component Gaussian mixture, NINT costs, K=20",
                   "n": n, "k": K, "nslots": int(g.nslots), "sweeps": int(st["sweeps"]),
                   "schedule": "reference", "node_order": "morton (f2m_spatial_order)",
                   "recovery": "exact where the reference recovery is not", "parallelism": "single GPU"},
        "solves_per_s": round(1e3 / ms, 3), "objective": st["objective"], "exact_ran": int(st["exact_ran"]),
        "order_ms": round(order_ms, 3), "knn_build_ms": round(knn_ms, 3), "input_order": input_order,
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "fwd_bwd": fwd_bwd,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def host_cpu():
    """nproc, the cores this process may use, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model}


def time_input_order(ctx, x, y, K, mode, args):
    """The same solve with nodes in input order (for comparison)."""
    g = ctx.knn_build(x, y, K, mode)
    for _ in range(args.warmup):
        ctx.solve(g)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.solve(g)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log("input order", ms, "ms/solve")
    return {"ms": round(ms, 3), "objective": st["objective"], "sweeps": int(st["sweeps"])}


def bench_batched(ctx, order, table=None, tune_to=None):
    import numpy as np

    from f2m_amd import workloads as W
    log("batched: generating 4096 instances")
    insts, mode, K = W.c4()
    xs = [a for a, _ in insts]
    ys = [b for _, b in insts]
    if order == "morton":
        perms = ctx.spatial_order_batched(xs, ys)
        xs = [a[p] for a, p in zip(xs, perms)]
        ys = [b[p] for b, p in zip(ys, perms)]
    t = time.perf_counter()
    b = ctx.knn_build_batched(xs, ys, K, mode)
    knn_s = time.perf_counter() - t
    tuned = None
    if tune_to:
        tune_batch_table(ctx, b, tune_to)
        table = tune_to
    if table and table != "none" and os.path.exists(table):
        ctx.batch_autotune(b, cache_path=table)  # loads the table
        tuned = {"file": os.path.relpath(table, ROOT),
                 "classes": [ln.split() for ln in open(table).read().splitlines() if ln.strip()]}
    t = time.perf_counter()
    st = ctx.solve_batched(b)  # first call captures the hipGraph
    first_s = time.perf_counter() - t
    t = time.perf_counter()
    st = ctx.solve_batched(b)  # replay
    batched_s = time.perf_counter() - t
    log("batched solve", batched_s, "s (capture call", first_s, "s)")
    idx = np.linspace(0, len(xs) - 1, 16).astype(int)
    loop_s = 0.0
    same = True
    for i in idx:
        g = ctx.knn_build(xs[i], ys[i], K, mode)
        ctx.solve(g)
        t = time.perf_counter()
        s2 = ctx.solve(g)
        loop_s += time.perf_counter() - t
        same &= s2["sweeps"] == st[i]["sweeps"] and abs(s2["objective"] - st[i]["objective"]) <= 1e-12 * abs(
            s2["objective"])
    plain_ms = loop_s * 1e3 / len(idx)
    per_inst_ms = batched_s * 1e3 / len(xs)
    return {"workload": "4096 instances, n log-uniform in [100,20000], fp32 points x1e6, NINT costs, K=16",
            "node_order": order, "instances": len(xs), "sum_n": int(sum(len(a) for a in xs)), "nslots": int(b.nslots),
            "ms": round(batched_s * 1e3, 1), "per_instance_ms": round(per_inst_ms, 4),
            "plain_loop_per_instance_ms": round(plain_ms, 3),
            "speedup_per_instance": round(plain_ms / per_inst_ms, 2), "identical_optima_sampled": bool(same),
            "all_optimal_found": all(x["status"] == 0 for x in st), "knn_build_s": round(knn_s, 3),
            "sweeps_max": int(max(x["sweeps"] for x in st)),
            "schedule_table": tuned or "built-in (lanes 2/4, 256/512/1024 threads from 1000/2000 nodes)"}


def tune_batch_table(ctx, b, path):
    """Schedule table of the batched leg, written to `path`: per instance-size
    class (n <= 500, 1000, 2000, 4000, 8000, 16000, larger) the lanes per node,
    the workgroup threads (over the classes: the bucket thresholds) and, for
    the classes above 8000 nodes, the shared-memory placement.  Scored by the
    makespan of the whole captured launch of the batch (the better of two
    timed replays after the capture call).  One pass of coordinate descent
    from the built-in schedule, classes from the largest down, a change kept
    when it is more than 0.5% faster."""
    B = [500, 1000, 2000, 4000, 8000, 16000, 2**63 - 1]
    # the built-in schedule (lanes 2 below 1000 nodes, else 4; 256 threads
    # below 1000, 512 below 2000, else 1024; default placement)
    tab = [[500, 2, 256, -1], [1000, 2, 256, -1], [2000, 4, 512, -1], [4000, 4, 1024, -1], [8000, 4, 1024, -1],
           [16000, 4, 1024, -1], [B[-1], 4, 1024, -1]]

    def write(t):
        with open(path, "w") as f:
            f.write("".join(" ".join(str(v) for v in r) + "\n" for r in t))

    def makespan(t):
        write(t)
        ctx.batch_autotune(b, cache_path=path)  # loads the table
        ctx.solve_batched(b)  # captures the graph
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            ctx.solve_batched(b)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    best = makespan(tab)
    log("batched: built-in table", round(best, 4), "s")
    cands = {B[6]: [(2, 1024, -1), (8, 1024, -1), (4, 1024, 2)], B[5]: [(2, 1024, -1), (8, 1024, -1), (4, 1024, 2)],
             B[4]: [(2, 1024, -1), (8, 1024, -1)], B[3]: [(2, 1024, -1), (8, 1024, -1)],
             B[2]: [(4, 1024, -1), (8, 512, -1), (2, 512, -1)], B[1]: [(4, 256, -1), (2, 512, -1)],
             B[0]: [(1, 256, -1), (4, 256, -1)]}
    for k in range(6, -1, -1):
        for l, bs, lds in cands[B[k]]:
            t = [list(r) for r in tab]
            t[k][1:] = [l, bs, lds if lds != -1 else t[k][3]]
            if t == tab:
                continue
            ms = makespan(t)
            log("batched: class", B[k], "lanes", l, "threads", bs, "placement", t[k][3], round(ms, 4), "s")
            if ms < best * 0.995:
                tab, best = t, ms
    write(tab)
    log("batched: chosen", tab, round(best, 4), "s ->", path)


def bench_fwd_bwd():
    import numpy as np
    import torch

    from f2m_amd import workloads as W
    from f2m_amd.autograd import F2MBound
    from f2m_amd.solver import Context
    insts, mode, K = W.c5()
    xs = [a for a, _ in insts]
    ys = [b for _, b in insts]
    node_off = np.zeros(len(xs) + 1, np.int64)
    node_off[1:] = np.cumsum([len(a) for a in xs])
    solver = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    px = torch.tensor(np.concatenate(xs), dtype=torch.float32, device="cuda", requires_grad=True)
    py = torch.tensor(np.concatenate(ys), dtype=torch.float32, device="cuda", requires_grad=True)
    best = None
    for rep in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        V = F2MBound.apply(px, py, node_off, K, solver, None)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        V.sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        px.grad = None
        py.grad = None
        if rep and (best is None or t2 - t0 < best[0] + best[1]):
            best = (t1 - t0, t2 - t1)
    log("fwd+bwd", best)
    # a training-loop step: points moved by ~1e-4, forward warm-started from
    # the previous step's potentials (WarmStart)
    from f2m_amd.autograd import WarmStart
    warm = WarmStart()
    F2MBound.apply(px.detach(), py.detach(), node_off, K, solver, None, warm)
    rng = np.random.default_rng(0)

    def moved(ax):
        return ax + 1e-4 * torch.tensor(rng.standard_normal(ax.numel()), dtype=torch.float32, device="cuda")

    # one untimed step first: the solver keeps the warm and the cold plan of a
    # batch (two captured graphs), so a training loop's steps replay both
    qx, qy = moved(px.detach()), moved(py.detach())
    F2MBound.apply(qx, qy, node_off, K, solver, None, warm)
    qx, qy = moved(qx), moved(qy)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Vw = F2MBound.apply(qx, qy, node_off, K, solver, None, warm)
    torch.cuda.synchronize()
    warm_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    Vc = F2MBound.apply(qx, qy, node_off, K, solver, None)
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    same = bool(torch.allclose(Vw, Vc, rtol=1e-9, atol=0))
    log("training step: warm", warm_ms, "cold", cold_ms, "same optima", same)
    solver.close()
    return {"workload": "256 instances of n=2000, fp32 points in [0,1)^2, RAW_FP32 costs, K=20",
            "fwd_ms": round(best[0] * 1e3, 2), "bwd_ms": round(best[1] * 1e3, 3),
            "fwd_bwd_ms": round((best[0] + best[1]) * 1e3, 2),
            "fwd_bwd_over_fwd": round((best[0] + best[1]) / best[0], 4),
            "step_fwd_warm_ms": round(warm_ms, 2), "step_fwd_cold_ms": round(cold_ms, 2),
            "step_same_optima": same}


if __name__ == "__main__":
    main()
