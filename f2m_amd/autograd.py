"""Differentiable F2M bound: torch.autograd.Function over the C ABI.

    V = F2MBound.apply(px, py, node_off, k, solver, opts, warm)

px, py: float32 CUDA tensors holding the points of a batch of instances
packed back to back (instance i = rows node_off[i]:node_off[i+1]).  Forward
builds the symmetrised kNN graph of each instance with RAW_FP32 costs
(float-rounded Euclidean distances), solves all instances with one batched
persistent launch and returns the optimum V_i of each instance (float64).
Backward runs the hand-written kernel f2m_bound_backward:
    dV_i/dp_u = sum_{e=(u,w)} x*_e (p_u - p_w) / |p_u - p_w|
(envelope theorem; valid where x* is unique and the kNN graph is locally
constant), scaled by the upstream gradient of the instance.

Each forward keeps its own copy of the graph (node offsets, row pointers
and neighbour ids, on the device) for its backward: the solver context's
batch buffers are overwritten by the next forward on the same context, so
several forwards may be pending before their backwards, in any order.

warm (optional, a WarmStart): keeps each forward's potentials and starts the
next forward of the same instances from them (f2m_solve_batched_warm), which
cuts the sweeps of a training loop whose points move a little per step.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .solver import COST_RAW_FP32, Context


class WarmStart:
    """Potentials of the last forward (device tensor), reused by the next
    forward when the batch has the same number of points."""

    def __init__(self):
        self.F = None


class F2MBound(torch.autograd.Function):
    @staticmethod
    def forward(ctx, px, py, node_off, k, solver: Context, opts=None, warm: WarmStart | None = None):
        if not (px.is_cuda and py.is_cuda and px.dtype == torch.float32 and py.dtype == torch.float32):
            raise TypeError("px, py must be float32 CUDA tensors")
        node_off = np.ascontiguousarray(np.asarray(node_off), np.int64)
        x64 = px.detach().to(torch.float64).contiguous()
        y64 = py.detach().to(torch.float64).contiguous()
        torch.cuda.current_stream().synchronize()
        b = solver.knn_build_batched((x64, node_off), y64, k, COST_RAW_FP32)
        F0 = warm.F if warm is not None and warm.F is not None and warm.F.numel() == b.n_total else None
        solver.solve_batched(b, opts, F_init=F0)
        xs = torch.empty(b.nslots, dtype=torch.float64, device=px.device)
        Fd = torch.empty(b.n_total, dtype=torch.float64, device=px.device)
        res = solver.batch_result(b, F=Fd, x=xs)
        if warm is not None:
            warm.F = Fd
        # this forward's graph, copied out of the context's reusable buffers
        dev = px.device
        g_off = torch.empty(b.n_inst + 1, dtype=torch.int64, device=dev)
        g_rp = torch.empty(b.n_total + 1, dtype=torch.int64, device=dev)
        g_nbr = torch.empty(b.nslots, dtype=torch.int32, device=dev)
        solver.batch_csr(b, out=(g_off, g_rp, g_nbr, None))
        ctx.save_for_backward(px.detach().contiguous(), py.detach().contiguous(), xs, g_off, g_rp, g_nbr)
        ctx.solver = solver
        ctx.shape = (b.n_inst, b.k, b.cost_mode, b.n_total, b.nslots)
        return torch.from_numpy(res["objective"]).to(px.device)

    @staticmethod
    def backward(ctx, gout):
        px, py, xs, g_off, g_rp, g_nbr = ctx.saved_tensors
        n_inst, k, mode, n_total, nslots = ctx.shape
        graph = L.Batch(n_inst, k, mode, 0, n_total, nslots, g_off.data_ptr(), g_rp.data_ptr(), g_nbr.data_ptr(),
                        None)
        scale = gout.detach().to(torch.float32).contiguous()
        gx = torch.empty_like(px)
        gy = torch.empty_like(py)
        torch.cuda.current_stream().synchronize()
        ctx.solver.bound_backward(graph, xs, px, py, scale, gx, gy)
        ctx.solver.synchronize()
        return gx, gy, None, None, None, None, None


def f2m_bound(px, py, node_off, k=20, solver: Context | None = None, opts=None, warm: WarmStart | None = None):
    """Convenience wrapper; creates a context on px's device when none is given."""
    solver = solver or Context(px.device.index or 0)
    return F2MBound.apply(px, py, node_off, k, solver, opts, warm)
