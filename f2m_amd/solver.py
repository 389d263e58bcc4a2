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

    def close(self):
        if self.h:
            lib.f2m_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        check(st, self.h)

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
    def solve(self, g: L.Graph | None = None, opts: L.Opts | None = None) -> dict:
        g = g or self.graph
        st = L.Stats()
        self._check(lib.f2m_solve(self.h, ctypes.byref(g), ctypes.byref(opts) if opts else None, ctypes.byref(st)))
        return st.as_dict()

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
