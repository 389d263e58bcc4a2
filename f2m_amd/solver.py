"""Host-side mirror of the reference interface over the C ABI.

The reference's seam is  graph(file, K) -> graph_c(graph*) -> graph::RUN()
(RUN.cpp:24-26); here it is  Context.knn_build(points, K) -> Context.solve(g).
Arrays may be numpy (host; staged by the library) or torch CUDA tensors
(device; passed by pointer).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L
from ._lib import check, lib

COST_RAW_FP64 = 0
COST_NINT = 1
COST_RAW_FP32 = 2


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def default_opts(**kw) -> L.Opts:
    o = L.Opts()
    lib.f2m_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


class Points:
    """TSPLIB points after 1e-3 rounding and duplicate removal (FXY.hpp:231-309)."""

    def __init__(self, path: str):
        h = ctypes.c_void_p()
        st = lib.f2m_tsplib_read(os.fsencode(path), ctypes.byref(h))
        check(st)
        self._h = h
        n = lib.f2m_points_count(h)
        self.labels = np.zeros(n, np.int64)
        self.x = np.zeros(n)
        self.y = np.zeros(n)
        lib.f2m_points_copy(h, _ptr(self.labels), _ptr(self.x), _ptr(self.y))
        self.n_duplicates = int(lib.f2m_points_duplicates(h))

    def write_duplicates(self, path: str):
        check(lib.f2m_points_write_duplicates(self._h, os.fsencode(path)))

    def __len__(self):
        return len(self.labels)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.f2m_points_free(self._h)
            self._h = None


def read_tsplib(path: str) -> Points:
    return Points(path)


class Context:
    """One f2m context: device workspace + a HIP stream (default: its own)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = ctypes.c_void_p()
        check(lib.f2m_ctx_create(device, stream, ctypes.byref(h)))
        self.h = h
        self.graph = None
        self.batch = None

    def close(self):
        if self.h:
            lib.f2m_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        check(st, self.h)

    def set_profile(self, sample_every: int):
        self._check(lib.f2m_ctx_set_profile(self.h, sample_every))

    def profile(self) -> dict:
        p = L.Profile()
        self._check(lib.f2m_get_profile(self.h, ctypes.byref(p)))
        return p.as_dict()

    def synchronize(self):
        self._check(lib.f2m_ctx_synchronize(self.h))

    # ---- init ------------------------------------------------------------
    def knn_directed(self, x, y, k: int, mode: int):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        kk = min(k, len(x) - 1)
        out = np.zeros((len(x), kk), np.int32)
        self._check(lib.f2m_knn_directed(self.h, len(x), _ptr(x), _ptr(y), k, mode, _ptr(out)))
        return out

    def knn_build(self, x, y, k: int, mode: int) -> L.Graph:
        if not hasattr(x, "data_ptr"):
            x = np.ascontiguousarray(x, np.float64)
            y = np.ascontiguousarray(y, np.float64)
        g = L.Graph()
        self._check(lib.f2m_knn_build(self.h, len(x), _ptr(x), _ptr(y), k, mode, ctypes.byref(g)))
        self.graph = g
        return g

    def spatial_order(self, x, y) -> np.ndarray:
        """Morton (Z-order) permutation of the points (f2m_spatial_order):
        perm[i] = input index of node i when the graph is built on
        x[perm], y[perm]."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        perm = np.zeros(len(x), np.int32)
        self._check(lib.f2m_spatial_order(self.h, len(x), _ptr(x), _ptr(y), _ptr(perm)))
        return perm

    def spatial_order_batched(self, xs, ys) -> list:
        """Per-instance Morton permutations (local indices) of a list of
        point sets, computed in one pass (f2m_spatial_order_batched)."""
        off = np.zeros(len(xs) + 1, np.int64)
        off[1:] = np.cumsum([len(a) for a in xs])
        x = np.ascontiguousarray(np.concatenate(xs), np.float64)
        y = np.ascontiguousarray(np.concatenate(ys), np.float64)
        perm = np.zeros(off[-1], np.int64)
        self._check(lib.f2m_spatial_order_batched(self.h, len(xs), _ptr(off), _ptr(x), _ptr(y), _ptr(perm)))
        return [perm[off[i]:off[i + 1]] - off[i] for i in range(len(xs))]

    def graph_from_csr(self, rowptr, nbr, cost) -> L.Graph:
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        cost = np.ascontiguousarray(cost, np.float64)
        g = L.Graph()
        self._check(lib.f2m_graph_from_csr(self.h, len(rowptr) - 1, _ptr(rowptr), _ptr(nbr), _ptr(cost),
                                           len(nbr), ctypes.byref(g)))
        self.graph = g
        return g

    def graph_csr(self, g: L.Graph | None = None):
        g = g or self.graph
        rowptr = np.zeros(g.n + 1, np.int32)
        nbr = np.zeros(g.nslots, np.int32)
        cost = np.zeros(g.nslots, np.float64)
        self._check(lib.f2m_graph_copy(self.h, ctypes.byref(g), _ptr(rowptr), _ptr(nbr), _ptr(cost)))
        return rowptr, nbr, cost

    # ---- solve -----------------------------------------------------------
    def solve(self, g: L.Graph | None = None, opts: L.Opts | None = None, F_init=None) -> dict:
        """f2m_solve, or f2m_solve_warm: the same schedule from potentials
        F_init (normalised, as result()["F"] returns them) instead of zero."""
        g = g or self.graph
        st = L.Stats()
        o = ctypes.byref(opts) if opts else None
        if F_init is None:
            self._check(lib.f2m_solve(self.h, ctypes.byref(g), o, ctypes.byref(st)))
        else:
            if not hasattr(F_init, "data_ptr"):
                F_init = np.ascontiguousarray(F_init, np.float64)
            self._check(lib.f2m_solve_warm(self.h, ctypes.byref(g), o, _ptr(F_init), ctypes.byref(st)))
        return st.as_dict()

    def knn_certify(self, g: L.Graph | None = None) -> int:
        """Nodes whose k-th nearest cost is below -d_min F after the last
        solve of g (f2m_knn_certify); 0 certifies that the optimum on the kNN
        graph is the LP optimum over all pairs of points."""
        g = g or self.graph
        v = ctypes.c_int64(0)
        self._check(lib.f2m_knn_certify(self.h, ctypes.byref(g), ctypes.byref(v)))
        return int(v.value)

    def solve_adaptive(self, x, y, k0: int, mode: int, k_max: int = 80, opts: L.Opts | None = None):
        """Adaptive K: solve with k0, and while the kNN restriction is not
        certified (knn_certify > 0) double k (at most k_max) and re-solve,
        warm-started from the previous potentials (node ids are unchanged).
        Returns (k, stats, result, history) with history = [(k, objective,
        violations, sweeps)]."""
        k = k0
        F = None
        hist = []
        while True:
            g = self.knn_build(x, y, k, mode)
            st = self.solve(g, opts, F_init=F)
            res = self.result()
            viol = self.knn_certify(g)
            hist.append((k, st["objective"], viol, st["sweeps"]))
            if viol == 0 or k >= k_max or k >= len(x) - 1:
                return k, st, res, hist
            F = res["F"]
            k = min(2 * k, k_max, len(x) - 1)

    def result(self, n: int | None = None, nslots: int | None = None):
        g = self.graph
        n = n or g.n
        nslots = nslots or g.nslots
        F = np.zeros(n)
        x = np.zeros(nslots)
        root = np.zeros(n, np.int32)
        self._check(lib.f2m_result_copy(self.h, _ptr(F), _ptr(x), _ptr(root)))
        return {"F": F, "x": x, "comb_root": root}

    def sweep(self, rowptr, nbr, D, length, F, nsweeps: int = 1):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        D = np.ascontiguousarray(D, np.float64)
        length = np.ascontiguousarray(length, np.int32)
        F = np.ascontiguousarray(F, np.float64)
        n = len(rowptr) - 1
        Fo = np.zeros(n)
        g = np.zeros(n)
        mx = ctypes.c_double(0)
        self._check(lib.f2m_sweep(self.h, n, _ptr(rowptr), _ptr(nbr), _ptr(D), _ptr(length), _ptr(F), _ptr(Fo),
                                  _ptr(g), nsweeps, ctypes.byref(mx)))
        return Fo, g, mx.value

    # ---- batched form ----------------------------------------------------
    def knn_build_batched(self, xs, ys, k: int, mode: int) -> L.Batch:
        """xs, ys: lists of per-instance coordinate arrays (or concatenated arrays
        with `node_off` given as xs=(x, node_off))."""
        if isinstance(xs, tuple):
            x, node_off = xs
            y = ys
        else:
            node_off = np.zeros(len(xs) + 1, np.int64)
            node_off[1:] = np.cumsum([len(a) for a in xs])
            x = np.concatenate(xs)
            y = np.concatenate(ys)
        if not hasattr(x, "data_ptr"):
            x = np.ascontiguousarray(x, np.float64)
            y = np.ascontiguousarray(y, np.float64)
        node_off = np.ascontiguousarray(node_off, np.int64)
        b = L.Batch()
        self._check(lib.f2m_knn_build_batched(self.h, len(node_off) - 1, _ptr(node_off), _ptr(x), _ptr(y), k, mode,
                                              ctypes.byref(b)))
        self.batch = b
        self.batch_node_off = node_off
        return b

    def batch_csr(self, b: L.Batch | None = None, out=None):
        """node_off, rowptr, nbr, cost of a batch; `out` = four destination
        arrays (numpy or CUDA tensors; None skips that array)."""
        b = b or self.batch
        if out is not None:
            self._check(lib.f2m_batch_copy(self.h, ctypes.byref(b), *(_ptr(a) for a in out)))
            return out
        node_off = np.zeros(b.n_inst + 1, np.int64)
        rowptr = np.zeros(b.n_total + 1, np.int64)
        nbr = np.zeros(b.nslots, np.int32)
        cost = np.zeros(b.nslots, np.float64)
        self._check(lib.f2m_batch_copy(self.h, ctypes.byref(b), _ptr(node_off), _ptr(rowptr), _ptr(nbr), _ptr(cost)))
        return node_off, rowptr, nbr, cost

    def solve_batched(self, b: L.Batch | None = None, opts: L.Opts | None = None, F_init=None) -> list:
        """f2m_solve_batched, or f2m_solve_batched_warm from potentials
        F_init (n_total, normalised; numpy or a CUDA tensor)."""
        b = b or self.batch
        st = (L.Stats * b.n_inst)()
        o = ctypes.byref(opts) if opts else None
        if F_init is None:
            self._check(lib.f2m_solve_batched(self.h, ctypes.byref(b), o, st))
        else:
            if not hasattr(F_init, "data_ptr"):
                F_init = np.ascontiguousarray(F_init, np.float64)
            self._check(lib.f2m_solve_batched_warm(self.h, ctypes.byref(b), o, _ptr(F_init), st))
        return [s.as_dict() for s in st]

    def batch_autotune(self, b: L.Batch | None = None, opts: L.Opts | None = None, sample: int = 256,
                       cache_path: str | None = None):
        b = b or self.batch
        self._check(lib.f2m_batch_autotune(self.h, ctypes.byref(b), ctypes.byref(opts) if opts else None, sample,
                                           os.fsencode(cache_path) if cache_path else None))

    def batch_result(self, b: L.Batch | None = None, F=None, x=None, objective=None, comb_root=False):
        """Copies results; pass torch tensors to keep them on the device.
        comb_root=True adds the per-node tight-component roots (local ids)."""
        b = b or self.batch
        F = np.zeros(b.n_total) if F is None else F
        x = np.zeros(b.nslots) if x is None else x
        objective = np.zeros(b.n_inst) if objective is None else objective
        self._check(lib.f2m_batch_result_copy(self.h, _ptr(F), _ptr(x), _ptr(objective)))
        out = {"F": F, "x": x, "objective": objective}
        if comb_root:
            root = np.zeros(b.n_total, np.int32)
            self._check(lib.f2m_batch_comb_root_copy(self.h, _ptr(root)))
            out["comb_root"] = root
        return out

    def bound_backward(self, b, x_slot, px, py, grad_scale, gx, gy):
        """Device tensors; b=None uses the context's single graph."""
        self._check(lib.f2m_bound_backward(self.h, ctypes.byref(b) if b is not None else None, _ptr(x_slot),
                                           _ptr(px), _ptr(py), _ptr(grad_scale), _ptr(gx), _ptr(gy)))

