"""ctypes binding of libdbie.so (the C ABI declared in include/dbie.h).

The library is built in-tree by ``dielectricbie.jl_amd/build.py``.  There is no
CPU fallback: if the library is missing, or no gfx950 GPU is visible when a
device entry point is called, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

# DBIE_LIB selects another build of the library (e.g. a sanitizer build of the host code)
LIB_PATH = os.environ.get("DBIE_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdbie.so")

OK, BAD_ARG, DIM_MISMATCH, OOM, HIP_ERROR, NOT_CONVERGED, NO_DEVICE, PENDING = range(8)
FP64, FP32, BF16, BF16X3, FP16X3 = 0, 1, 2, 3, 4
OP_DT, OP_S, OP_D = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "bf16x3": BF16X3, "fp16x3": FP16X3}


class DbieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[dbie status {code}] {msg}")
        self.code = code


class NotConverged(DbieError):
    pass


class NearOpts(C.Structure):
    _fields_ = [("up_tol", C.c_double), ("max_order", C.c_int), ("correct_edges", C.c_int),
                ("adaptive_atol", C.c_double), ("adaptive_rtol", C.c_double), ("adaptive_n_GL", C.c_int),
                ("adaptive_max_depth", C.c_int), ("build_corrections", C.c_int), ("n_threads", C.c_int),
                ("device_build", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("itmax", C.c_int), ("restart", C.c_int),
                ("precondition", C.c_int), ("precision", C.c_int), ("inner_rtol", C.c_double),
                ("max_refine", C.c_int), ("check_residual", C.c_int), ("krylov", C.c_int),
                ("recycle", C.c_int), ("verify", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("niter", C.c_int), ("nrefine", C.c_int), ("solved", C.c_int), ("max_rel_residual", C.c_double),
                ("time_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_double), ("setup_workers", C.c_int), ("setup_threads", C.c_int),
                ("solve_slots", C.c_int), ("max_batch_rhs", C.c_int), ("cache_ops", C.c_int),
                ("inject_oom_every", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("state", C.c_int), ("N", C.c_int64), ("K", C.c_int),
                ("batch_K", C.c_int), ("solved", C.c_int), ("niter", C.c_int), ("nrefine", C.c_int),
                ("max_rel_residual", C.c_double), ("op_cache_hit", C.c_int), ("queue_ms", C.c_double),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double), ("total_ms", C.c_double),
                ("message", C.c_char * 256)]

    STATES = ("queued", "setup", "ready", "solving", "done", "failed", "cancelled")

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["message"] = self.message.decode(errors="replace")
        d["state"] = self.STATES[self.state] if 0 <= self.state < len(self.STATES) else str(self.state)
        return d


class ServerSummary(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("completed", "failed", "operator_builds", "cache_hits", "batches",
                                         "fused_requests", "cached_operators", "bytes_in_use", "peak_bytes",
                                         "budget_bytes", "multi_geometry_batches")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> argtypes; every one returns dbie_status unless listed in _RESTYPES
SIGNATURES = {
    "dbie_last_error": [C.c_char_p, C.c_size_t],
    "dbie_init": [C.c_int],
    "dbie_near_opts_default": [C.POINTER(NearOpts)],
    "dbie_solve_opts_default": [C.POINTER(SolveOpts)],
    "dbie_gauss_legendre": [C.c_int, _dp, _dp],
    "dbie_best_grid": [C.c_double, C.c_double, C.c_double, _ip, _ip],
    "dbie_geometry_box": [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(_vp)],
    "dbie_geometry_from_panels": [C.c_int64, C.c_int, _dp, _dp, _ip, C.POINTER(_vp)],
    "dbie_geometry_from_panel_nodes": [C.c_int64, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp, C.POINTER(_vp)],
    "dbie_geometry_multi_box": [C.c_int, C.c_double, C.c_int, _dp, _dp, C.c_double, C.c_double, C.POINTER(_vp)],
    "dbie_geom_eps": [_vp, _dp, _dp],
    "dbie_geom_sizes": [_vp, _i64p, _i64p, _ip],
    "dbie_geom_nodes": [_vp, _dp],
    "dbie_geom_panels": [_vp, _dp, _dp, _ip],
    "dbie_geom_destroy": [_vp],
    "dbie_neighbor_list": [_vp, C.c_int, C.c_double, C.c_int, _i64p, _i64p, _ip, _ip, _ip, _ip, _ip],
    "dbie_operator_create": [_vp, _dp, _dp, C.POINTER(NearOpts), C.POINTER(_vp)],
    "dbie_operator_destroy": [_vp],
    "dbie_op_sizes": [_vp, _i64p, _i64p, _ip],
    "dbie_op_set_eps": [_vp, _dp, _dp],
    "dbie_op_corrections": [_vp, _ip, _ip, _dp],
    "dbie_op_corrections_kernel": [_vp, C.c_int, _ip, _ip, _dp],
    "dbie_op_device_bytes": [_vp, _i64p],
    "dbie_op_setup_ms": [_vp, _dp],
    "dbie_apply": [_vp, C.c_int, _dp, C.c_int64, _dp, C.c_int64, C.c_int],
    "dbie_apply_kernel": [_vp, C.c_int, C.c_int, _dp, C.c_int64, _dp, C.c_int64],
    "dbie_apply_kernel_corrected": [_vp, C.c_int, C.c_int, _dp, C.c_int64, _dp, C.c_int64],
    "dbie_lhs_dense": [_vp, _dp, C.c_int64],
    "dbie_solve_lu": [C.c_int64, _dp, C.c_int64, C.c_int, _dp, C.c_int64],
    "dbie_rhs_points": [_vp, C.c_int, _dp, _dp, C.c_double, _dp, C.c_int64],
    "dbie_rhs_volume": [_vp, C.c_int, _i64p, _dp, _dp, _dp, C.c_int, C.c_double, _dp, C.c_int64],
    "dbie_rhs_volume_boxes": [_vp, C.c_int, _i64p, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_int,
                              C.c_double, _dp, C.c_int64],
    "dbie_screen_boxes": [C.c_int64, _dp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double,
                          C.c_int, _dp],
    "dbie_screen_interface": [_vp, C.c_int64, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp],
    "dbie_solve_block": [_vp, C.c_int, _dp, C.c_int64, _dp, C.c_int64, C.POINTER(SolveOpts), C.POINTER(SolveStats)],
    "dbie_pottrg": [_vp, C.c_int64, _dp, C.c_int, _dp, C.c_int64, _dp, C.c_int64],
    "dbie_op_stream": [_vp],
    "dbie_op_synchronize": [_vp],
    "dbie_solve_fused": [C.c_int, C.POINTER(_vp), _ip, C.POINTER(_dp), _i64p, C.POINTER(_dp), _i64p,
                         C.POINTER(SolveOpts), C.POINTER(SolveStats)],
    "dbie_op_residual_history": [_vp, C.c_int, C.POINTER(C.c_int), _dp],
    "dbie_autotune": [C.c_int, C.c_int],
    "dbie_autotune_choice": [_vp, C.c_int, C.c_int, _ip, _ip],
    "dbie_autotune_set": [_vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "dbie_op_graphs": [_vp, C.c_int, _i64p, _i64p],
    "dbie_op_stored_far": [_vp, C.c_int, _i64p],
    "dbie_op_profile": [_vp, C.c_int],
    "dbie_op_profile_read": [_vp, C.c_int, _i64p, _dp],
    "dbie_apply_device": [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, C.c_int],
    "dbie_apply_device_noalloc": [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, C.c_int],
    "dbie_solve_block_device": [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, C.POINTER(SolveOpts),
                                C.POINTER(SolveStats)],
    "dbie_device_alloc": [C.c_size_t, C.POINTER(_vp)],
    "dbie_device_free": [_vp],
    "dbie_memcpy_h2d": [_vp, _vp, C.c_size_t],
    "dbie_memcpy_d2h": [_vp, _vp, C.c_size_t],
    "dbie_copy_block_device": [C.c_int64, C.c_int, _vp, C.c_int64, _vp, C.c_int64],
    "dbie_version": [],
    "dbie_tmap_create": [_vp, C.c_int64, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)],
    "dbie_tmap_create_ex": [_vp, C.c_int64, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(_vp)],
    "dbie_tmap_setup_ms": [_vp, _dp],
    "dbie_tmap_destroy": [_vp],
    "dbie_tmap_sizes": [_vp, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p],
    "dbie_tmap_refined": [_vp, _dp, _ip, _ip],
    "dbie_tmap_matrices": [_vp, _i64p, _ip, _dp, _i64p, _ip, _dp],
    "dbie_tmap_apply": [_vp, C.c_int, _dp, C.c_int64, _dp, C.c_int64],
    "dbie_tmap_apply_device": [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64],
    "dbie_reaction_matrix": [_vp, C.c_int, _i64p, _dp, C.c_int, _dp, C.c_int64, _dp, C.c_int64],
    "dbie_reaction_matrix_device": [_vp, C.c_int, _i64p, _dp, C.c_int, _vp, C.c_int64, _dp, C.c_int64],
    "dbie_gauss_hermite": [C.c_int, _dp, _dp],
    "dbie_server_opts_default": [C.POINTER(ServerOpts)],
    "dbie_server_create": [C.POINTER(ServerOpts), C.POINTER(_vp)],
    "dbie_server_destroy": [_vp],
    "dbie_server_submit_box": [_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                               C.c_double, C.c_int, _dp, _dp, C.POINTER(SolveOpts), _i64p],
    "dbie_server_wait": [_vp, C.c_int64, C.c_double, C.POINTER(RequestInfo)],
    "dbie_server_take": [_vp, C.c_int64, _dp, C.c_int64],
    "dbie_server_cancel": [_vp, C.c_int64],
    "dbie_server_pause": [_vp, C.c_int],
    "dbie_server_stats": [_vp, C.POINTER(ServerSummary)],
}
_RESTYPES = {"dbie_version": C.c_char_p, "dbie_op_stream": _vp, "dbie_geom_destroy": None,
             "dbie_operator_destroy": None, "dbie_tmap_destroy": None, "dbie_server_destroy": None, "dbie_server_opts_default": None, "dbie_near_opts_default": None, "dbie_solve_opts_default": None}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libdbie.so not found at {LIB_PATH}; build it with "
                              "`python dielectricbie.jl_amd/build.py` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().dbie_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status: int):
    if status != OK:
        msg = last_error()
        if status == NOT_CONVERGED:
            raise NotConverged(status, msg)
        if status == BAD_ARG:
            raise ValueError(msg)
        if status == DIM_MISMATCH:
            raise ValueError("DimensionMismatch: " + msg)
        raise DbieError(status, msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64
    return a.ctypes.data_as(_i64p)


def fortran(a, rows: int | None = None) -> np.ndarray:
    """Column-major float64 copy/view of a vector or matrix (Julia memory order)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if rows is not None and a.shape[0] != rows:
        raise ValueError(f"DimensionMismatch: expected {rows} rows, got {a.shape[0]}")
    return np.asfortranarray(a)
