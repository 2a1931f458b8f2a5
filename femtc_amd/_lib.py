"""Loader for libfemtc.so, the HIP (gfx950) library behind the C ABI in
include/femtc.h.  There is no CPU fallback: if the library is missing this
module raises, so a GPU run can never silently compute elsewhere."""
from __future__ import annotations

import ctypes as C
import os
from functools import lru_cache

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfemtc.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_DEGREE_NOT_SUPPORTED = 2
ERR_ORDER_NOT_SUPPORTED = 3
ERR_BLOCK_NOT_SUPPORTED = 4
ERR_INVALID_ORDER_STRING = 5
ERR_HIP = 6
ERR_OUT_OF_MEMORY = 7
ERR_NOT_CONVERGED = 8
ERR_BREAKDOWN = 9
ERR_NO_DEVICE = 10

F64, F32, BF16, BF16_ROUNDED = 0, 1, 2, 3
MODE_CORRECT, MODE_REFERENCE = 0, 1
OP_MASS, OP_STIFFNESS, OP_KPM = 0, 1, 2

_vp = C.c_void_p
_i32 = C.c_int
_i64 = C.c_int64
_sz = C.c_size_t
_dp = C.POINTER(C.c_double)

# name -> (restype, argtypes); the exported surface of include/femtc.h
SIGNATURES = {
    "femtc_status_string": (C.c_char_p, [_i32]),
    "femtc_last_error": (C.c_char_p, []),
    "femtc_version": (_i32, []),
    "femtc_max_degree": (_i32, []),
    "femtc_max_fused_degree": (_i32, []),
    "femtc_string2order": (_i32, [C.c_char_p, C.POINTER(_i32)]),
    "femtc_check_ladder": (_i32, [_i32, _i32, _i32, _i32]),
    "femtc_device_count": (_i32, [C.POINTER(_i32)]),
    "femtc_device_alloc": (_i32, [C.POINTER(_vp), _sz]),
    "femtc_device_free": (_i32, [_vp]),
    "femtc_memcpy_htod": (_i32, [_vp, _vp, _sz]),
    "femtc_memcpy_dtoh": (_i32, [_vp, _vp, _sz]),
    "femtc_device_synchronize": (_i32, []),
    "femtc_gll_points": (_i32, [_i32, _dp]),
    "femtc_gauss_rule": (_i32, [_i32, _dp, _dp]),
    "femtc_basis_tables": (_i32, [_i32, _i32, _dp]),
    "femtc_mass_apply": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp,
                                _i64, _vp]),
    "femtc_stiffness_apply": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp,
                                     _vp, _i64, _vp]),
    "femtc_kpm_apply": (_i32, [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                               _i64, _vp]),
    "femtc_global_apply": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _i64,
                                  _vp, _vp, _vp, _i64, _vp]),
    "femtc_global_apply_add": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp,
                                      _i64, _vp, _vp, _vp, _i64, _vp]),
    "femtc_fill_zero": (_i32, [_vp, _i64, _i32, _vp]),
    "femtc_stream_probe": (_i32, [_i32, _i32, _vp, _vp, _i64, _vp, _i32,
                                  _i32, _vp]),
    "femtc_set_high_degree_kernel": (_i32, [_i32, C.POINTER(_i32)]),
    "femtc_apply_kernel": (_i32, [_i32, _i32, _i32, _i32, _i32,
                                  C.POINTER(_i32)]),
    "femtc_mass_operator": (_i32, [_i32, _vp, _sz, _vp, _vp, _vp, _sz, _i32,
                                   _i32, _i32, _i32, _i32]),
    "femtc_stiffness_operator": (_i32, [_i32, _vp, _sz, _vp, _vp, _vp, _i32,
                                        _i32, _i32, _i32, _i32, _i32]),
    "femtc_dofmap_structured": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp]),
    "femtc_geometry_affine": (_i32, [_i64, _i32, _vp, _vp, _vp, _i32, _vp]),
    "femtc_geometry_cell": (_i32, [_i64, _vp, _vp, _i32, _vp]),
    "femtc_basis_tables_weights": (_i32, [_i32, _i32, _dp]),
    "femtc_even_odd_tables": (_i32, [_i32, _i32, _dp, _dp, _dp,
                                     C.POINTER(_i64), C.POINTER(_i32)]),
    "femtc_global_apply_affine": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp,
                                         _i64, _vp, _vp, _i64, _i32, _vp]),
    "femtc_mfma_tables_bytes": (_i32, [_i32, C.POINTER(_sz)]),
    "femtc_mfma_tables": (_i32, [_i32, _i32, _i32, _dp, _vp]),
    "femtc_mfma_tables_host": (_i32, [_i32, _i32, _i32, _dp, _vp]),
    "femtc_global_apply_mfma": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp,
                                       _i64, _i32, _i32, _i32, _vp, _vp, _i32,
                                       _vp]),
    "femtc_global_apply_mfma_noalloc": (_i32, [_i32, _i32, _i32, _i32, _vp,
                                               _vp, _vp, _i64, _i32, _i32,
                                               _i32, _vp, _i32, _vp]),
    "femtc_plan_create": (_i32, [C.POINTER(_vp), _i32, _i32, _i32, _i32, _vp,
                                 _i64, _vp]),
    "femtc_plan_destroy": (_i32, [_vp]),
    "femtc_plan_set_block": (_i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32]),
    "femtc_plan_get_block": (_i32, [_vp, _i32, _i32, _i32, C.POINTER(_i32),
                                    C.POINTER(_i32), C.POINTER(_i32)]),
    "femtc_plan_apply": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                _i32, _vp]),
    "femtc_plan_autotune": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp,
                                   _vp, _i32, _vp, C.POINTER(C.c_double)]),
    "femtc_plan_set_tables": (_i32, [_vp, _i32, _i32, _vp, _vp,
                                     C.POINTER(_i32)]),
    "femtc_batch_create": (_i32, [C.POINTER(_vp), _vp, _i32, _vp]),
    "femtc_batch_run": (_i32, [_vp]),
    "femtc_batch_graph_launch": (_i32, [_vp]),
    "femtc_batch_autotune": (_i32, [_vp, _i32]),
    "femtc_batch_stats": (_i32, [_vp, C.POINTER(_i32), C.POINTER(_i64),
                                 C.POINTER(_i64)]),
    "femtc_batch_group": (_i32, [_vp, _i32] + [C.POINTER(_i32)] * 7),
    "femtc_batch_set_group": (_i32, [_vp, _i32, _i32]),
    "femtc_batch_destroy": (_i32, [_vp]),
    "femtc_server_create": (_i32, [C.POINTER(_vp), _i32, _sz]),
    "femtc_server_register": (_i32, [_vp, _vp]),
    "femtc_server_unregister": (_i32, [_vp, _vp]),
    "femtc_server_submit": (_i32, [_vp, _vp]),
    "femtc_server_drain": (_i32, [_vp]),
    "femtc_server_results": (_i32, [_vp, _vp, _i32, C.POINTER(_i32)]),
    "femtc_server_memory": (_i32, [_vp, C.POINTER(_sz), C.POINTER(_i32)]),
    "femtc_server_destroy": (_i32, [_vp]),
    "femtc_global_diagonal": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _i64,
                                     _vp, _vp, _vp, _i64, _vp]),
    "femtc_solver_create": (_i32, [C.POINTER(_vp), _vp, _i32, _vp]),
    "femtc_solver_solve": (_i32, [_vp, _vp, _vp, C.c_double, C.c_double, _i32,
                                  _vp]),
    "femtc_solver_diagonal": (_i32, [_vp, C.POINTER(_vp)]),
    "femtc_solver_scatter": (_i32, [_vp, C.POINTER(_i32)]),
    "femtc_solver_even_odd": (_i32, [_vp, C.POINTER(_i32)]),
    "femtc_solver_set_batch": (_i32, [_vp, _i32, _i32]),
    "femtc_solver_destroy": (_i32, [_vp]),
    "femtc_solver_bytes": (_i32, [_vp, C.POINTER(_sz)]),
    "femtc_solver_workspace_bytes": (_i32, [_vp, _i32, C.POINTER(_sz)]),
    "femtc_admission_simulate": (_i32, [_vp, _i32, _i32, _vp, _i32, _sz, _vp,
                                        _vp]),
}


class Problem(C.Structure):
    """femtc_problem (include/femtc.h)."""
    _fields_ = [("op", _i32), ("degree", _i32), ("nq", _i32),
                ("num_cells", _i64), ("ndofs", _i64), ("dofmap", _vp),
                ("tables", _vp), ("G", _vp), ("detJw", _vp), ("J", _vp),
                ("grid", _i32 * 3), ("scatter", _i32), ("even_odd", _i32)]


SCATTER_AUTO, SCATTER_DIRECT, SCATTER_ASSEMBLED, SCATTER_ZRUN = 0, 1, 2, 3

HIGH_DEGREE_AUTO, HIGH_DEGREE_GENERAL = 0, 1
KERNEL_FUSED, KERNEL_GENERAL, KERNEL_LINE = 0, 1, 2


class BatchItem(C.Structure):
    """femtc_batch_item (include/femtc.h)."""
    _fields_ = [("op", _i32), ("dtype", _i32), ("degree", _i32), ("nq", _i32),
                ("tables", _vp), ("x", _vp), ("y", _vp), ("G", _vp),
                ("detJw", _vp), ("dofmap", _vp), ("num_cells", _i64),
                ("ndofs", _i64), ("grid", _i32 * 3)]


REQUEST_APPLY, REQUEST_SOLVE = 0, 1


class Request(C.Structure):
    """femtc_request (include/femtc.h)."""
    _fields_ = [("kind", _i32), ("problem", _vp), ("input", _vp),
                ("output", _vp), ("rtol", C.c_double),
                ("inner_rtol", C.c_double), ("maxit", _i32),
                ("inner_dtype", _i32), ("tag", _i64)]


class RequestResult(C.Structure):
    """femtc_request_result (include/femtc.h)."""
    _fields_ = [("tag", _i64), ("kind", _i32), ("status", _i32),
                ("slot", _i32), ("iterations", _i32),
                ("outer_iterations", _i32), ("relres", C.c_double),
                ("t_arrival", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double)]


class SimRequest(C.Structure):
    """femtc_sim_request (include/femtc.h)."""
    _fields_ = [("t_arrival", C.c_double), ("kind", _i32), ("problem", _i32),
                ("inner", _i32), ("bytes", _sz), ("actual", _sz),
                ("service", C.c_double), ("create", C.c_double)]


class SimResult(C.Structure):
    """femtc_sim_result (include/femtc.h)."""
    _fields_ = [("status", _i32), ("slot", _i32), ("cached", _i32),
                ("t_start", C.c_double), ("t_end", C.c_double)]


class SimStats(C.Structure):
    """femtc_sim_stats (include/femtc.h)."""
    _fields_ = [("high_water", _sz), ("evictions", _i64), ("created", _i64),
                ("destroyed", _i64), ("makespan", C.c_double),
                ("left_in_queue", _i64)]


class SolveInfo(C.Structure):
    """femtc_solve_info (include/femtc.h)."""
    _fields_ = [("status", _i32), ("outer_iterations", _i32),
                ("inner_iterations", _i32), ("relres", C.c_double),
                ("seconds", C.c_double)]


class FemtcError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: {status_string(status)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


@lru_cache(maxsize=None)
def lib() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m femtc_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def status_string(status: int) -> str:
    return lib().femtc_status_string(status).decode()


def last_error() -> str:
    return lib().femtc_last_error().decode()


def check(status: int, where: str) -> None:
    if status != OK:
        raise FemtcError(status, where, last_error())
