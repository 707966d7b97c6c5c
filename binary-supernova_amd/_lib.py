"""ctypes binding of libbsn_hip.so (include/bsn.h).

The library is the product: nothing here falls back to a CPU path.  If the
shared object is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsn_hip.so")

BSN_OK = 0
BSN_HOST = 0
BSN_DEVICE = 1
PREC = {"fp64": 0, "split": 1}

# every symbol include/bsn.h declares
EXPORTS = [
    "bsn_version", "bsn_last_error", "bsn_ctx_create", "bsn_ctx_destroy", "bsn_ctx_synchronize",
    "bsn_ctx_info", "bsn_grid_create", "bsn_grid_destroy", "bsn_grid_set_state",
    "bsn_grid_get_state", "bsn_grid_set_phi", "bsn_grid_set_params", "bsn_grid_device_state",
    "bsn_fill_ghost_3d", "bsn_apply_floors_3d", "bsn_euler3d_rhs", "bsn_euler3d_step",
    "bsn_cfl_dt_3d", "bsn_apply_sg_force", "bsn_rk3_save", "bsn_rk3_combine",
    "bsn_conserved_totals", "bsn_add_bh_gravity_source", "bsn_add_sink_sources",
    "bsn_grid_set_point_masses", "bsn_grid_set_sinks", "bsn_solve_poisson_isolated",
    "bsn_grid_set_self_gravity", "bsn_gas_force_on_bh", "bsn_accrete", "bsn_gas_diagnostics",
    "bsn_grid_export_slabs", "bsn_grid_import_slabs", "bsn_grid_dims",
    "bsn_euler3d_step_host", "bsn_cfl_dt_3d_host", "bsn_batch_create",
    "bsn_batch_destroy", "bsn_batch_autotune", "bsn_batch_set_shapes", "bsn_batch_set_mode",
    "bsn_batch_set_update_shape", "bsn_batch_step",
    "bsn_batch_run", "bsn_batch_times", "bsn_fmr_create", "bsn_fmr_destroy",
    "bsn_fmr_prolong_all", "bsn_fmr_prolong_fine_ghosts", "bsn_fmr_restrict", "bsn_fmr3d_step",
    "bsn_cfl_dt_fmr3d", "bsn_fmr_run", "bsn_fmr_time", "bsn_fmr_set_mode", "bsn_fmr_autotune", "bsn_batch_profile",
    "bsn_fmr3d_step_gravity", "bsn_fmr_set_self_gravity", "bsn_solve_fmr_poisson",
    "bsn_fmr_restrict_density", "bsn_fmr_prolong_scalar", "bsn_conserved_totals_noalloc",
    "bsn_batch_get_shapes",
]


class BsnError(RuntimeError):
    pass


_lib = None
_d, _i, _i64, _p = C.c_double, C.c_int, C.c_int64, C.c_void_p
_pd = C.POINTER(C.c_double)


# int (*)(void* user, int part, int64_t k0, int64_t nk, double* data, int64_t count)
SLAB_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int64,
                      C.POINTER(C.c_double), C.c_int64)
LAYOUT_FIELDS = 0
LAYOUT_REF = 1


def load():
    """Load libbsn_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BsnError(f"{LIB_PATH} is missing: build it with binary-supernova_amd/build.py")
    L = C.CDLL(LIB_PATH)
    sig = {
        "bsn_version": ([], _i),
        "bsn_last_error": ([], C.c_char_p),
        "bsn_ctx_create": ([_i, C.POINTER(_p)], _i),
        "bsn_ctx_destroy": ([_p], _i),
        "bsn_ctx_synchronize": ([_p], _i),
        "bsn_ctx_info": ([_p, C.c_char_p, _i], _i),
        "bsn_grid_create": ([_p, _i64, _i64, _i64, _i, C.POINTER(_p)], _i),
        "bsn_grid_destroy": ([_p], _i),
        "bsn_grid_set_state": ([_p, _p, _i], _i),
        "bsn_grid_get_state": ([_p, _p, _i], _i),
        "bsn_grid_set_phi": ([_p, _p, _i], _i),
        "bsn_grid_set_params": ([_p, _d, _d, _d, _d, _i, _d, _d], _i),
        "bsn_grid_device_state": ([_p, C.POINTER(_p), C.POINTER(_i64), C.POINTER(_i64)], _i),
        "bsn_fill_ghost_3d": ([_p, _i], _i),
        "bsn_apply_floors_3d": ([_p, _d, _d, _d], _i),
        "bsn_euler3d_rhs": ([_p, _p, _p, _p, _p, _i, _d, _d, _d, _d, _i, _d, _d, _i], _i),
        "bsn_euler3d_step": ([_p, _d, _d, _d, _d, _d, _i, _d, _d], _i),
        "bsn_cfl_dt_3d": ([_p, _d, _d, _d, _d, _d, _pd], _i),
        "bsn_apply_sg_force": ([_p, _p, _i, _d, _d, _d], _i),
        "bsn_rk3_save": ([_p], _i),
        "bsn_rk3_combine": ([_p, _i, _p, _i, _d], _i),
        "bsn_conserved_totals": ([_p, _d, _pd], _i),
        "bsn_conserved_totals_noalloc": ([_p, _d, _p, _i], _i),
        "bsn_add_bh_gravity_source": ([_p, _p, _i, _d, _d, _d, _i, _pd, _pd, _pd, _d, _d, _d], _i),
        "bsn_add_sink_sources": ([_p, _p, _i, _d, _d, _d, _i, _pd, _pd, _pd, _pd, _i, _d, _d, _d, _d],
                                 _i),
        "bsn_grid_set_point_masses": ([_p, _i, _pd, _pd, _pd, _d, _d, _d], _i),
        "bsn_grid_set_sinks": ([_p, _i, _pd, _pd, _pd, _pd, _i, _d, _d, _d, _d], _i),
        "bsn_solve_poisson_isolated": ([_p, _p, _i, _p, _i, _d, _d, _d], _i),
        "bsn_grid_set_self_gravity": ([_p, _i], _i),
        "bsn_gas_force_on_bh": ([_p, _d, _d, _d, _i, _p, _p, _p, _d, _d, _d, _p], _i),
        "bsn_accrete": ([_p, _d, _d, _d, _i, _p, _p, _p, _p, _p, _d, _d, _d, _d, _d, _p], _i),
        "bsn_gas_diagnostics": ([_p, _d, _d, _d, _d, _d, _d, _i, _p, _p, _p, _d, _p], _i),
        "bsn_grid_export_slabs": ([_p, _i, _i64, SLAB_FN, _p], _i),
        "bsn_grid_import_slabs": ([_p, _i, _i64, SLAB_FN, _p], _i),
        "bsn_grid_dims": ([_p, C.POINTER(_i64)], _i),
        "bsn_euler3d_step_host": ([_p, _i64, _i64, _i64, _d, _d, _d, _d, _d, _i, _d, _d, _i], _i),
        "bsn_cfl_dt_3d_host": ([_p, _i64, _i64, _i64, _d, _d, _d, _d, _d, _pd], _i),
        "bsn_batch_create": ([_p, _i, C.POINTER(_p), C.POINTER(_p)], _i),
        "bsn_batch_destroy": ([_p], _i),
        "bsn_batch_autotune": ([_p, C.c_char_p, _i], _i),
        "bsn_batch_set_shapes": ([_p, _i, _i, _i], _i),
        "bsn_batch_set_mode": ([_p, _i], _i),
        "bsn_batch_set_update_shape": ([_p, _i], _i),
        "bsn_batch_get_shapes": ([_p, C.POINTER(_i)], _i),
        "bsn_batch_step": ([_p, _pd], _i),
        "bsn_batch_run": ([_p, _d, _d, _i, _i], _i),
        "bsn_batch_times": ([_p, _pd, _pd], _i),
        "bsn_batch_profile": ([_p, _i, _pd, _pd, _pd], _i),
        "bsn_fmr_create": ([_p, _i, C.POINTER(_p), C.POINTER(_i64), C.POINTER(_i), C.POINTER(_p)], _i),
        "bsn_fmr_destroy": ([_p], _i),
        "bsn_fmr_prolong_all": ([_p, _i], _i),
        "bsn_fmr_prolong_fine_ghosts": ([_p, _i], _i),
        "bsn_fmr_restrict": ([_p, _i], _i),
        "bsn_fmr3d_step": ([_p, _d], _i),
        "bsn_cfl_dt_fmr3d": ([_p, _d, _pd], _i),
        "bsn_fmr_run": ([_p, _d, _d, _i, _i], _i),
        "bsn_fmr_time": ([_p, _pd, _pd], _i),
        "bsn_fmr_set_mode": ([_p, _i], _i),
        "bsn_fmr3d_step_gravity": ([_p, _d, _i, _d], _i),
        "bsn_fmr_set_self_gravity": ([_p, _i, _d], _i),
        "bsn_solve_fmr_poisson": ([_p, C.POINTER(_p), _i], _i),
        "bsn_fmr_restrict_density": ([_p, _i], _i),
        "bsn_fmr_prolong_scalar": ([_p, _i, _p, _i, _p, _i], _i),
        "bsn_fmr_autotune": ([_p, C.c_char_p, _i], _i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != BSN_OK:
        msg = load().bsn_last_error().decode(errors="replace")
        raise BsnError(f"{what}: status {rc}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
