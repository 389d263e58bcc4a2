"""ctypes binding of libf2m_hip.so (the C ABI in include/f2m.h).

There is no fallback: if the HIP library is missing, importing this module
raises ImportError.  Build it with `python -m f2m_amd.build`.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libf2m_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"f2m_amd: HIP library not built ({LIB_PATH}); run `python -m f2m_amd.build`")

# One HIP runtime per process: PyTorch ships its own libamdhip64 with the same
# soname as /opt/rocm's.  Loading torch first makes the dynamic loader bind this
# library to torch's runtime, so tensors and f2m calls share device state.
try:  # pragma: no cover - torch is optional
    import torch  # noqa: F401
except ImportError:
    pass

lib = ctypes.CDLL(LIB_PATH)

c_i32, c_i64, c_f64, c_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class Opts(ctypes.Structure):
    _fields_ = [
        ("err_grad", c_f64), ("err_comb_mult", c_f64), ("glob_k", c_i32), ("glob_rounds", c_i32),
        ("glob_sweeps", c_i32), ("glob_exit", c_f64), ("max_outer", c_i32), ("factor0", c_i32),
        ("factor_step", c_i32), ("factor_max", c_i32), ("check_every", c_i32), ("early_exit", c_i32),
        ("repair", c_i32), ("shift0", c_f64), ("shift_max", c_f64), ("fix_rounds", c_i32),
        ("fix_factor", c_i32), ("frozen_sweeps", c_i32), ("stop_phase", c_i32),
        ("exact", c_i32), ("exact_alpha", c_i32), ("exact_eps0", c_f64), ("exact_tol", c_f64),
        ("exact_max_rounds", c_i64), ("exact_max_phases", c_i32), ("exact_reserved", c_i32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("objective", c_f64), ("dual", c_f64), ("max_grad", c_f64), ("d_min", c_f64), ("sweeps", c_i64),
        ("glob_rounds", c_i32), ("outer_rounds", c_i32), ("n_combs", c_i32), ("n_irregular_a", c_i32),
        ("n_irregular_b", c_i32), ("n_repairs", c_i32), ("status", c_i32), ("max_rank", c_i32),
        ("exact_ran", c_i32), ("exact_phases", c_i32), ("exact_rounds", c_i64), ("exact_gap", c_f64),
        ("exact_certified", c_i32), ("exact_reserved", c_i32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Graph(ctypes.Structure):
    _fields_ = [
        ("n", c_i32), ("k", c_i32), ("cost_mode", c_i32), ("reserved", c_i32), ("nslots", c_i64),
        ("d_min", c_f64), ("rowptr", c_p), ("nbr", c_p), ("cost", c_p),
    ]


class Batch(ctypes.Structure):
    _fields_ = [
        ("n_inst", c_i32), ("k", c_i32), ("cost_mode", c_i32), ("reserved", c_i32), ("n_total", c_i64),
        ("nslots", c_i64), ("node_off", c_p), ("rowptr", c_p), ("nbr", c_p), ("cost", c_p),
    ]


class Profile(ctypes.Structure):
    _fields_ = [
        ("sweep_launches", c_i64), ("sweep_samples", c_i64), ("sweep_ms_sampled", c_f64),
        ("sweep_alg_bytes", c_f64), ("active_slots", c_i64), ("solve_ms", c_f64),
        ("sweeps", c_i64), ("sampled_sweeps", c_i64), ("sampled_alg_bytes", c_f64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol declared in include/f2m.h
EXPORTS = [
    "f2m_default_opts", "f2m_strerror", "f2m_last_error", "f2m_version", "f2m_ctx_create", "f2m_ctx_destroy",
    "f2m_ctx_synchronize", "f2m_ctx_set_profile", "f2m_get_profile", "f2m_tsplib_read", "f2m_points_count", "f2m_points_duplicates", "f2m_points_copy",
    "f2m_points_write_duplicates", "f2m_points_free", "f2m_knn_build", "f2m_knn_directed", "f2m_graph_from_csr",
    "f2m_graph_copy", "f2m_solve", "f2m_result_copy", "f2m_sweep", "f2m_knn_build_batched", "f2m_solve_batched", "f2m_batch_autotune", "f2m_batch_copy",
    "f2m_batch_result_copy", "f2m_bound_backward", "f2m_spatial_order", "f2m_spatial_order_batched",
    "f2m_solve_warm", "f2m_knn_certify", "f2m_solve_batched_warm", "f2m_batch_comb_root_copy",
    "f2m_sweep_noalloc",
]

_sig = {
    "f2m_default_opts": (None, [ctypes.POINTER(Opts)]),
    "f2m_strerror": (ctypes.c_char_p, [c_i32]),
    "f2m_last_error": (ctypes.c_char_p, [c_p]),
    "f2m_version": (ctypes.c_char_p, []),
    "f2m_ctx_create": (c_i32, [c_i32, c_p, ctypes.POINTER(c_p)]),
    "f2m_ctx_destroy": (None, [c_p]),
    "f2m_ctx_synchronize": (c_i32, [c_p]),
    "f2m_ctx_set_profile": (c_i32, [c_p, c_i32]),
    "f2m_get_profile": (c_i32, [c_p, ctypes.POINTER(Profile)]),
    "f2m_tsplib_read": (c_i32, [ctypes.c_char_p, ctypes.POINTER(c_p)]),
    "f2m_points_count": (c_i64, [c_p]),
    "f2m_points_duplicates": (c_i64, [c_p]),
    "f2m_points_copy": (None, [c_p, c_p, c_p, c_p]),
    "f2m_points_write_duplicates": (c_i32, [c_p, ctypes.c_char_p]),
    "f2m_points_free": (None, [c_p]),
    "f2m_knn_build": (c_i32, [c_p, c_i32, c_p, c_p, c_i32, c_i32, ctypes.POINTER(Graph)]),
    "f2m_knn_directed": (c_i32, [c_p, c_i32, c_p, c_p, c_i32, c_i32, c_p]),
    "f2m_graph_from_csr": (c_i32, [c_p, c_i32, c_p, c_p, c_p, c_i64, ctypes.POINTER(Graph)]),
    "f2m_graph_copy": (c_i32, [c_p, ctypes.POINTER(Graph), c_p, c_p, c_p]),
    "f2m_solve": (c_i32, [c_p, ctypes.POINTER(Graph), ctypes.POINTER(Opts), ctypes.POINTER(Stats)]),
    "f2m_result_copy": (c_i32, [c_p, c_p, c_p, c_p]),
    "f2m_sweep": (c_i32, [c_p, c_i32, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i32, c_p]),
    "f2m_knn_build_batched": (c_i32, [c_p, c_i32, c_p, c_p, c_p, c_i32, c_i32, ctypes.POINTER(Batch)]),
    "f2m_solve_batched": (c_i32, [c_p, ctypes.POINTER(Batch), ctypes.POINTER(Opts), c_p]),
    "f2m_batch_autotune": (c_i32, [c_p, ctypes.POINTER(Batch), ctypes.POINTER(Opts), c_i32, ctypes.c_char_p]),
    "f2m_batch_copy": (c_i32, [c_p, ctypes.POINTER(Batch), c_p, c_p, c_p, c_p]),
    "f2m_batch_result_copy": (c_i32, [c_p, c_p, c_p, c_p]),
    "f2m_bound_backward": (c_i32, [c_p, ctypes.POINTER(Batch), c_p, c_p, c_p, c_p, c_p, c_p]),
    "f2m_spatial_order": (c_i32, [c_p, c_i32, c_p, c_p, c_p]),
    "f2m_spatial_order_batched": (c_i32, [c_p, c_i32, c_p, c_p, c_p, c_p]),
    "f2m_solve_warm": (c_i32, [c_p, ctypes.POINTER(Graph), ctypes.POINTER(Opts), c_p, ctypes.POINTER(Stats)]),
    "f2m_knn_certify": (c_i32, [c_p, ctypes.POINTER(Graph), c_p]),
    "f2m_solve_batched_warm": (c_i32, [c_p, ctypes.POINTER(Batch), ctypes.POINTER(Opts), c_p, c_p]),
    "f2m_batch_comb_root_copy": (c_i32, [c_p, c_p]),
    "f2m_sweep_noalloc": (c_i32, [c_p, c_i32, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i32, c_p]),
}

for _name in EXPORTS:
    _f = getattr(lib, _name, None)
    if _f is not None and _name in _sig:
        _f.restype, _f.argtypes = _sig[_name]


class F2MError(RuntimeError):
    pass


def check(status: int, ctx=None):
    if status != 0:
        msg = lib.f2m_strerror(status).decode()
        if ctx is not None:
            detail = lib.f2m_last_error(ctx).decode()
            if detail:
                msg += ": " + detail
        raise F2MError(f"f2m status {status}: {msg}")
