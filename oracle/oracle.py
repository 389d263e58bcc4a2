"""ctypes wrapper of the CPU oracle (oracle/libf2m_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The f2m_amd product never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libf2m_oracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        _setup(_lib)
    return _lib


class Opts(ctypes.Structure):
    _fields_ = [
        ("err_grad", ctypes.c_double),
        ("err_comb_mult", ctypes.c_double),
        ("glob_k", ctypes.c_int32),
        ("glob_rounds", ctypes.c_int32),
        ("glob_sweeps", ctypes.c_int32),
        ("glob_exit", ctypes.c_double),
        ("max_outer", ctypes.c_int32),
        ("factor0", ctypes.c_int32),
        ("factor_step", ctypes.c_int32),
        ("factor_max", ctypes.c_int32),
        ("check_every", ctypes.c_int32),
        ("early_exit", ctypes.c_int32),
        ("repair", ctypes.c_int32),
        ("shift0", ctypes.c_double),
        ("shift_max", ctypes.c_double),
        ("fix_rounds", ctypes.c_int32),
        ("fix_factor", ctypes.c_int32),
        ("frozen_sweeps", ctypes.c_int32),
        ("stop_phase", ctypes.c_int32),
        ("exact", ctypes.c_int32),
        ("exact_alpha", ctypes.c_int32),
        ("exact_eps0", ctypes.c_double),
        ("exact_tol", ctypes.c_double),
        ("exact_max_rounds", ctypes.c_int64),
        ("exact_max_phases", ctypes.c_int32),
        ("exact_reserved", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_double),
        ("dual", ctypes.c_double),
        ("max_grad", ctypes.c_double),
        ("d_min", ctypes.c_double),
        ("sweeps", ctypes.c_int64),
        ("glob_rounds", ctypes.c_int32),
        ("outer_rounds", ctypes.c_int32),
        ("n_combs", ctypes.c_int32),
        ("n_irregular_a", ctypes.c_int32),
        ("n_irregular_b", ctypes.c_int32),
        ("n_repairs", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("max_rank", ctypes.c_int32),
        ("exact_ran", ctypes.c_int32),
        ("exact_phases", ctypes.c_int32),
        ("exact_rounds", ctypes.c_int64),
        ("exact_gap", ctypes.c_double),
        ("exact_certified", ctypes.c_int32),
        ("exact_reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _setup(L):
    L.orc_default_opts.argtypes = [ctypes.POINTER(Opts)]
    L.orc_cost.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int]
    L.orc_cost.restype = ctypes.c_double
    L.orc_knn_directed.argtypes = [ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _P]
    L.orc_graph_build.argtypes = [ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]
    L.orc_graph_build.restype = _P
    L.orc_graph_nslots.argtypes = [_P]
    L.orc_graph_nslots.restype = ctypes.c_int64
    L.orc_graph_export.argtypes = [_P, _P, _P, _P]
    L.orc_graph_free.argtypes = [_P]
    L.orc_sweep.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    L.orc_qpart.argtypes = [_P, _P, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_qpart.restype = ctypes.c_int
    L.orc_solve.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.POINTER(Opts), ctypes.c_int, _P, _P, _P,
                            ctypes.POINTER(Stats)]
    L.orc_solve.restype = ctypes.c_int
    L.orc_axb.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_double)]
    L.orc_axb.restype = ctypes.c_int
    L.orc_bound_backward.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    L.orc_exact.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_double, _P, ctypes.POINTER(Opts), _P, _P, _P,
                            ctypes.POINTER(Stats)]
    L.orc_exact.restype = ctypes.c_int
    L.orc_solve_warm.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.POINTER(Opts), ctypes.c_int, _P, _P, _P, _P,
                                 ctypes.POINTER(Stats)]
    L.orc_solve_warm.restype = ctypes.c_int
    L.orc_knn_certify.argtypes = [ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_double, _P]
    L.orc_knn_certify.restype = ctypes.c_int64


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().orc_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def cost(xu, yu, xw, yw, mode):
    return lib().orc_cost(xu, yu, xw, yw, mode)


def knn_directed(x, y, K, mode, brute=False, threads=8):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = len(x)
    K = min(K, n - 1)
    out = np.zeros((n, K), np.int32)
    rc = lib().orc_knn_directed(n, _ptr(x), _ptr(y), K, mode, int(brute), threads, _ptr(out))
    assert rc == 0
    return out


def graph_build(x, y, K, mode, brute=False, threads=8):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = len(x)
    h = lib().orc_graph_build(n, _ptr(x), _ptr(y), K, mode, int(brute), threads)
    assert h
    m = lib().orc_graph_nslots(h)
    rowptr = np.zeros(n + 1, np.int32)
    nbr = np.zeros(m, np.int32)
    cst = np.zeros(m, np.float64)
    lib().orc_graph_export(h, _ptr(rowptr), _ptr(nbr), _ptr(cst))
    lib().orc_graph_free(h)
    return rowptr, nbr, cst


def sweep(rowptr, nbr, D, length, F):
    n = len(rowptr) - 1
    F = np.ascontiguousarray(F, np.float64)
    Fo = np.zeros(n)
    g = np.zeros(n)
    lib().orc_sweep(n, _ptr(np.ascontiguousarray(rowptr, np.int32)), _ptr(np.ascontiguousarray(nbr, np.int32)),
                    _ptr(np.ascontiguousarray(D, np.float64)), _ptr(np.ascontiguousarray(length, np.int32)),
                    _ptr(F), _ptr(Fo), _ptr(g))
    return Fo, g


def qpart(d, thr, raise_):
    d = np.array(d, np.float64)
    idx = np.arange(len(d), dtype=np.int32)
    p = lib().orc_qpart(_ptr(d), _ptr(idx), len(d), thr, int(raise_))
    return p, d, idx


def solve(rowptr, nbr, cst, opts: Opts | None = None, threads=8, F_init=None):
    n = len(rowptr) - 1
    m = len(nbr)
    o = opts or default_opts()
    F = np.zeros(n)
    x = np.zeros(m)
    root = np.zeros(n, np.int32)
    st = Stats()
    fi = None if F_init is None else np.ascontiguousarray(F_init, np.float64)
    rc = lib().orc_solve_warm(n, _ptr(np.ascontiguousarray(rowptr, np.int32)),
                              _ptr(np.ascontiguousarray(nbr, np.int32)), _ptr(np.ascontiguousarray(cst, np.float64)),
                              ctypes.byref(o), threads, _ptr(fi), _ptr(F), _ptr(x), _ptr(root), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_solve failed: {rc}")
    return {"F": F, "x": x, "comb_root": root, "stats": st.as_dict()}


def axb(rows, b, nv):
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    bb = np.array(b, np.int32)
    x = np.zeros(nv)
    err = ctypes.c_double(0)
    go = lib().orc_axb(len(rows), _ptr(rp), _ptr(col), _ptr(bb), nv, _ptr(x), ctypes.byref(err))
    return go, x, err.value


def bound_backward(rowptr, nbr, x_slot, px, py):
    n = len(rowptr) - 1
    gx = np.zeros(n)
    gy = np.zeros(n)
    lib().orc_bound_backward(n, _ptr(np.ascontiguousarray(rowptr, np.int32)), _ptr(np.ascontiguousarray(nbr, np.int32)),
                             _ptr(np.ascontiguousarray(x_slot, np.float64)), _ptr(np.ascontiguousarray(px, np.float64)),
                             _ptr(np.ascontiguousarray(py, np.float64)), _ptr(gx), _ptr(gy))
    return gx, gy


def exact(rowptr, nbr, cst, F, d_min=None, opts: Opts | None = None):
    """Exact recovery from potentials F (normalised units): returns rc and a
    dict with x per canonical slot, the certifying F, half_root and stats."""
    n = len(rowptr) - 1
    m = len(nbr)
    cst = np.ascontiguousarray(cst, np.float64)
    if d_min is None:
        d_min = float(cst[cst > 0].min())
    o = opts or default_opts()
    Fo = np.zeros(n)
    x = np.zeros(m)
    hr = np.zeros(n, np.int32)
    st = Stats()
    rc = lib().orc_exact(n, _ptr(np.ascontiguousarray(rowptr, np.int32)), _ptr(np.ascontiguousarray(nbr, np.int32)),
                         _ptr(cst), d_min, _ptr(np.ascontiguousarray(F, np.float64)), ctypes.byref(o), _ptr(Fo),
                         _ptr(x), _ptr(hr), ctypes.byref(st))
    return rc, {"F": Fo, "x": x, "half_root": hr, "stats": st.as_dict()}


def knn_certify(rowptr, cst, K, d_min, F):
    """Nodes whose K-th nearest cost is below -d_min F (0: the kNN graph's
    optimum is the complete graph's)."""
    n = len(rowptr) - 1
    return int(lib().orc_knn_certify(n, _ptr(np.ascontiguousarray(rowptr, np.int32)),
                                     _ptr(np.ascontiguousarray(cst, np.float64)), K, d_min,
                                     _ptr(np.ascontiguousarray(F, np.float64))))
