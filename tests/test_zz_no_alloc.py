"""GPU test of the allocation-free entry point f2m_sweep_noalloc.

Both entry points are called through the C ABI on the same device arrays:
f2m_sweep_noalloc must give f2m_sweep's potentials, gradient and max|g| bit
for bit, on two calls in a row, the second on a smaller graph (the fixed
scratch array keeps the first call's contents), with max|g| returned to host
memory and to device memory; arguments it must refuse are refused before
any launch.
"""
import ctypes

import numpy as np
import pytest

from f2m_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _case(ctx, n, seed, K):
    x, y = W.uniform_int(n, 10**6, seed)
    g = ctx.knn_build(x, y, K, W.COST_NINT)
    rp, nb, c = ctx.graph_csr(g)
    D = c / c[c > 0].min()
    deg = np.diff(rp)
    rng = np.random.default_rng(seed)
    length = np.minimum(deg, rng.integers(3, K + 6, size=len(deg))).astype(np.int32)
    F = rng.uniform(-3.0, 1.0, size=len(deg))
    return rp, nb, D, length, F


def _noalloc(lib, h, arrs, n, nsweeps, max_dev):
    import torch
    rp, nb, D, ln, F = arrs
    Fo = torch.empty(n, dtype=torch.float64, device="cuda")
    go = torch.empty(n, dtype=torch.float64, device="cuda")
    if max_dev:
        mx_t = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
        mx = mx_t.data_ptr()
    else:
        mx_h = ctypes.c_double(-1.0)
        mx = ctypes.addressof(mx_h)
    st = lib.f2m_sweep_noalloc(h, n, rp.data_ptr(), nb.data_ptr(), D.data_ptr(), ln.data_ptr(), F.data_ptr(),
                               Fo.data_ptr(), go.data_ptr(), nsweeps, mx)
    assert st == 0, lib.f2m_last_error(h)
    torch.cuda.synchronize()
    m = mx_t.item() if max_dev else mx_h.value
    return Fo.cpu().numpy(), go.cpu().numpy(), m


def test_sweep_noalloc_matches_sweep():
    import torch

    from f2m_amd._lib import lib
    from f2m_amd.solver import Context
    ctx = Context(0, stream=torch.cuda.current_stream().cuda_stream)
    for n, seed, K, nsweeps, max_dev in ((3000, 21, 16, 5, False), (700, 22, 10, 4, True), (500, 23, 8, 1, False)):
        rp, nb, D, ln, F = _case(ctx, n, seed, K)
        say("f2m_sweep", n, nsweeps)
        Fr, gr, mr = ctx.sweep(rp, nb, D, ln, F, nsweeps)
        dev = [torch.tensor(a, device="cuda") for a in (rp, nb, D, ln, F)]
        say("f2m_sweep_noalloc", n, nsweeps, "max to device" if max_dev else "max to host")
        Fo, go, mo = _noalloc(lib, ctx.h, dev, n, nsweeps, max_dev)
        assert np.array_equal(Fo, Fr) and np.array_equal(go, gr) and mo == mr, (n, nsweeps)
        assert mo == np.abs(gr).max()
    # refused before any launch: host arrays, aliased F, n above the limit
    dev = [torch.tensor(a, device="cuda") for a in (rp, nb, D, ln, F)]
    Fo = torch.empty(len(ln), dtype=torch.float64, device="cuda")
    hF = np.zeros(len(ln))
    st = lib.f2m_sweep_noalloc(ctx.h, len(ln), dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                               dev[3].data_ptr(), dev[4].data_ptr(), hF.ctypes.data, None, 1, None)
    assert st == -1 and b"device arrays" in lib.f2m_last_error(ctx.h)
    st = lib.f2m_sweep_noalloc(ctx.h, len(ln), dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                               dev[3].data_ptr(), dev[4].data_ptr(), dev[4].data_ptr(), None, 1, None)
    assert st == -1
    st = lib.f2m_sweep_noalloc(ctx.h, (1 << 20) + 1, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                               dev[3].data_ptr(), dev[4].data_ptr(), Fo.data_ptr(), None, 1, None)
    assert st == -6 and b"F2M_NOALLOC_MAX_N" in lib.f2m_last_error(ctx.h)
    ctx.close()
