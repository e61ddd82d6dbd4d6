"""ctypes binding of libpdds.so (include/pdds_engine.h).

The library is loaded from this package directory.  There is no fallback:
if the HIP library is missing, or no device is present, the calls fail.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PDDS_LIB") or os.path.join(_HERE, "libpdds.so")

OK, ERR_INVALID, ERR_NOMEM, ERR_DEVICE, ERR_UNSUPPORTED, ERR_SINGULAR, ERR_NOT_READY, ERR_FULL = range(8)
PDE_MONEGROS, PDE_SINSIN, PDE_VARCOEF, PDE_VARCOEF_LUT, PDE_HELMHOLTZ, PDE_TABLE = range(6)
STEP_CV, STEP_PLAIN, STEP_VR, STEP_VRCV = range(4)
FP64, FP32 = 0, 1

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class Problem(C.Structure):
    _fields_ = [("pde", C.c_int32), ("k2", C.c_double), ("lut", C.c_void_p), ("refl", C.c_int32),
                ("robin", C.c_double), ("rho", C.c_double), ("ulut", C.c_void_p), ("vlut", C.c_void_p)]


SIDE_N, SIDE_S, SIDE_E, SIDE_W = 1, 2, 4, 8


class Run(C.Structure):
    _fields_ = [("h", C.c_double), ("beta", C.c_double), ("seed", C.c_uint32), ("stream", C.c_uint32),
                ("precision", C.c_int32), ("quirk_last_side", C.c_int32), ("max_steps", C.c_int64),
                ("sub_batch", C.c_int32), ("item_paths", C.c_int32), ("wg_size", C.c_int32),
                ("grid_wgs", C.c_int32), ("lut_lds", C.c_int32), ("plain", C.c_int32),
                ("paths_per_lane", C.c_int32), ("step", C.c_int32), ("adaptive_h", C.c_int32),
                ("lut_pad", C.c_int32)]


class Adaptive(C.Structure):
    _fields_ = [("se_target", C.c_double), ("n_min", C.c_int64), ("n_max", C.c_int64), ("n_pilot", C.c_int64),
                ("max_rounds", C.c_int32), ("tol", C.c_double), ("growth", C.c_double), ("variance", C.c_int32)]


class AdaptiveReport(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("iters", C.c_int32), ("relres", C.c_double), ("paths", C.c_int64),
                ("path_steps", C.c_int64), ("max_se", C.c_double), ("knots_capped", C.c_int32),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("stream", C.c_uint32),
                ("knots_over", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneConfig(C.Structure):
    _fields_ = [("sample_paths", C.c_int64), ("max_knots", C.c_int32), ("reps", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("best_rate", C.c_double), ("base_rate", C.c_double),
                ("total_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Counters(C.Structure):
    _fields_ = [("path_steps", C.c_int64), ("paths", C.c_int64), ("exits", C.c_int64 * 6),
                ("capped", C.c_int64), ("kernel_ms", C.c_double), ("reduce_ms", C.c_double),
                ("total_ms", C.c_double), ("work_items", C.c_int64), ("grid_wgs", C.c_int32),
                ("wg_size", C.c_int32), ("item_paths", C.c_int32), ("sub_batch", C.c_int32)]

    def as_dict(self):
        return {"path_steps": self.path_steps, "paths": self.paths, "exits": list(self.exits),
                "capped": self.capped, "kernel_ms": self.kernel_ms, "reduce_ms": self.reduce_ms,
                "total_ms": self.total_ms, "work_items": self.work_items, "grid_wgs": self.grid_wgs,
                "wg_size": self.wg_size, "item_paths": self.item_paths, "sub_batch": self.sub_batch}


class ServerConfig(C.Structure):
    _fields_ = [("problem", Problem), ("run", Run), ("domain", C.c_double * 4), ("hbm_budget", C.c_int64),
                ("max_inflight", C.c_int32), ("prep_threads", C.c_int32), ("round_paths", C.c_int64)]


REQ_POINT, REQ_SOLVE = 0, 1


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("stream", C.c_uint32), ("k2", C.c_double), ("n_paths", C.c_int64),
                ("x", C.c_double), ("y", C.c_double), ("n_sub", C.c_int32), ("n_node", C.c_int32),
                ("fac", C.c_double * 4), ("tol", C.c_double), ("h", C.c_double)]


class Result(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("n_knots", C.c_int32), ("iters", C.c_int32),
                ("relres", C.c_double), ("u", C.c_double), ("u_var", C.c_double), ("path_steps", C.c_int64),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = {
    "pdds_engine_create": (C.c_int32, [C.c_int, C.POINTER(C.c_void_p)]),
    "pdds_engine_destroy": (None, [C.c_void_p]),
    "pdds_last_error": (C.c_char_p, [C.c_void_p]),
    "pdds_status_string": (C.c_char_p, [C.c_int32]),
    "pdds_engine_info": (C.c_int32, [C.c_void_p, _i32, C.c_char_p, C.c_int32]),
    "pdds_run_defaults": (None, [C.POINTER(Run)]),
    "pdds_lut_create": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _d, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_void_p)]),
    "pdds_lut_destroy": (None, [C.c_void_p]),
    "pdds_lut_eval": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, _d, _d, _d]),
    "pdds_mesh_create": (C.c_int32, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, _d,
                                     C.POINTER(C.c_void_p)]),
    "pdds_mesh_create_disc": (C.c_int32, [C.c_double, C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_int32,
                                          C.c_double, C.POINTER(C.c_void_p)]),
    "pdds_mesh_geometry": (C.c_int32, [C.c_void_p, C.c_int32, _i32, _d]),
    "pdds_mesh_shape": (C.c_int32, [C.c_void_p, C.c_int32, _d]),
    "pdds_mesh_entry_shapes": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _d, _i32]),
    "pdds_mesh_destroy": (None, [C.c_void_p]),
    "pdds_mesh_n_knots": (C.c_int32, [C.c_void_p]),
    "pdds_mesh_n_stencils": (C.c_int32, [C.c_void_p]),
    "pdds_mesh_knots": (C.c_int32, [C.c_void_p, _d, _d, _d, _d, _i32, _i32]),
    "pdds_mesh_stencil": (C.c_int32, [C.c_void_p, C.c_int32, _d, _i32, _i32, C.c_int32, _i32, _d, _d, _i32]),
    "pdds_mesh_cond": (C.c_int32, [C.c_void_p, C.c_int32, _d]),
    "pdds_mesh_csr": (C.c_int32, [C.c_void_p, _i64, _i32, _i64]),
    "pdds_mesh_ipsi": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, _d, _i32]),
    "pdds_paths": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(Run), C.c_double, C.c_double, _d,
                               C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _d, _d, _d, _d, _d, _i64,
                               _d, _d, _d, _d, _d, _i64, C.POINTER(Counters)]),
    "pdds_solve_knot": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run), C.c_int32,
                                    C.c_int64, _d, _d, _d, _d, C.POINTER(Counters)]),
    "pdds_knot_steps": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run), _d]),
    "pdds_assemble": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run), _i64, C.c_int64,
                                  _d, _d, _d, _d, C.POINTER(Counters)]),
    "pdds_assemble_eq": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run), _i64, C.c_int64,
                                     _d, _d, _d, _d, _d, _d, _d, C.POINTER(Counters)]),
    "pdds_adaptive_defaults": (None, [C.POINTER(Adaptive)]),
    "pdds_solve_adaptive": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run),
                                        C.POINTER(Adaptive), _d, _i64, _d, C.POINTER(AdaptiveReport)]),
    "pdds_autotune": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Problem), C.POINTER(Run), C.POINTER(TuneConfig),
                                  C.POINTER(Run), C.POINTER(TuneReport)]),
    "pdds_point_eval": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(Run), _d, C.c_int32, _d, _d, _i64,
                                    _d, _d, C.POINTER(Counters)]),
    "pdds_point_eval_noalloc": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(Run), _d, C.c_int32, _d, _d,
                                            _i64, _d, _d, C.POINTER(Counters)]),
    "pdds_solve": (C.c_int32, [C.c_void_p, C.c_int32, _i64, _i32, _d, _d, _d, C.c_double, C.c_int32, _i32, _d]),
    "pdds_server_create": (C.c_int32, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "pdds_server_destroy": (None, [C.c_void_p]),
    "pdds_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "pdds_server_poll": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(Result), _d, C.c_int64]),
    "pdds_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(Result), _d, C.c_int64]),
    "pdds_server_stats": (C.c_int32, [C.c_void_p, _i64, _i64, _i64, _i64, _d]),
}

_LIB = None


def lib():
    """Load libpdds.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpdds.so is not built (run __graft_entry__.build() or make -C pdds_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class PddsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (lib().pdds_status_string(status).decode(), msg))
        self.status = status


def ptr(a, t=_d):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
