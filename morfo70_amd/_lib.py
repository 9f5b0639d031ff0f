"""ctypes loader of libm70h.so (include/m70h.h). Fails loudly when the library is missing:
there is no CPU fallback for the Hydro step."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libm70h.so")

M70H_OK, M70H_EINVAL, M70H_EHIP, M70H_ENONFINITE, M70H_ESTATE = 0, 1, 2, 3, 4
SOLVER_EXACT = 0
SOLVER_PARTITION = 1
SOLVER_FUSED = 2

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None

# every symbol include/m70h.h declares, with (restype, argtypes)
SIGNATURES = {
    "m70h_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _vp,
                                   ctypes.POINTER(_vp)]),
    "m70h_destroy": (None, [_vp]),
    "m70h_last_error": (ctypes.c_char_p, [_vp]),
    "m70h_set_params": (ctypes.c_int, [_vp, _vp]),
    "m70h_set_stream": (ctypes.c_int, [_vp, _vp]),
    "m70h_init_from_bathymetry": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "m70h_set_state": (ctypes.c_int, [_vp] + [_dp] * 10),
    "m70h_get_state": (ctypes.c_int, [_vp] + [_dp] * 10),
    "m70h_set_state_dev": (ctypes.c_int, [_vp] + [_vp] * 10),
    "m70h_get_state_dev": (ctypes.c_int, [_vp] + [_vp] * 10),
    "m70h_set_forcing": (ctypes.c_int, [_vp] + [_dp] * 7),
    "m70h_set_forcing_dev": (ctypes.c_int, [_vp] + [_vp] * 7),
    "m70h_set_offshore_f0": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m70h_set_offshore_f0_series": (ctypes.c_int, [_vp, _dp, ctypes.c_long]),
    "m70h_bc_constants": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "m70h_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_run_async": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_synchronize": (ctypes.c_int, [_vp]),
    "m70h_set_solver": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_set_solver_tiles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "m70h_set_rows_block": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_tridiag_solve": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_long] + [_vp] * 5 +
                           [ctypes.c_long, _vp, ctypes.c_long, ctypes.c_int]),
    "m70h_tridiag_solve_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_long] + [_vp] * 5 +
                                   [ctypes.c_long, ctypes.c_int]),
    "m70h_check_finite": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "m70h_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_set_profiling_only": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "m70h_profile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp,
                                    ctypes.POINTER(ctypes.c_long)]),
    "m70h_state_ptr": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "m70h_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "m70h_batch_destroy": (None, [_vp]),
    "m70h_batch_last_error": (ctypes.c_char_p, [_vp]),
    "m70h_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_batch_set_rows_block": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_batch_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "m70h_batch_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_batch_run_async": (ctypes.c_int, [_vp, ctypes.c_int]),
    "m70h_batch_synchronize": (ctypes.c_int, [_vp]),
    "m70s_create": (ctypes.c_int, [_vp, _dp, _dp, _vp, ctypes.POINTER(_vp)]),
    "m70s_destroy": (None, [_vp]),
    "m70s_last_error": (ctypes.c_char_p, [_vp]),
    "m70s_calculate_fluxes": (ctypes.c_int, [_vp]),
    "m70s_update_bottom": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m70s_calculate_fluxes_async": (ctypes.c_int, [_vp]),
    "m70s_update_bottom_async": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m70s_set_fluxes": (ctypes.c_int, [_vp, _dp, _dp]),
    "m70s_get": (ctypes.c_int, [_vp] + [_dp] * 8),
    "m70s_set_bed": (ctypes.c_int, [_vp, _dp, _dp]),
    "m70w_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _vp, ctypes.POINTER(_vp)]),
    "m70w_destroy": (None, [_vp]),
    "m70w_last_error": (ctypes.c_char_p, [_vp]),
    "m70w_set_stream": (ctypes.c_int, [_vp, _vp]),
    "m70w_set_depth": (ctypes.c_int, [_vp, _dp]),
    "m70w_set_uv": (ctypes.c_int, [_vp, _dp, _dp]),
    "m70w_calculate": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "m70w_calculate_async": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "m70w_get": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "m70w_steps": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_long)]),
    "m70w_couple_hydro": (ctypes.c_int, [_vp, _vp]),
    "m70w_set_uv_from_hydro": (ctypes.c_int, [_vp]),
    "m70w_push_forcing": (ctypes.c_int, [_vp]),
    "m70o_create": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    "m70o_destroy": (None, [_vp]),
    "m70o_last_error": (ctypes.c_char_p, [_vp]),
    "m70o_run": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_long)]),
    "m70o_time": (ctypes.c_int, [_vp, _dp]),
    "m70o_records": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_long), _dp]),
    "m70o_snapshot_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_long)]),
    "m70o_snapshot": (ctypes.c_int, [_vp, ctypes.c_long, _dp, ctypes.c_char_p, _dp]),
    "m70nc_dims": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "m70nc_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_long]),
    "m70nc_read_att": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _dp]),
    "m70nc_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_dp),
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                   _dp]),
    "m70nc_last_error": (ctypes.c_char_p, []),
    "m70o_save_snapshot": (ctypes.c_int, [_vp, ctypes.c_long, ctypes.c_char_p, _dp, _dp, _dp]),
    "m70_resume": (ctypes.c_int, [ctypes.c_char_p, _vp, _vp, _vp, ctypes.c_int, _dp]),
    "m70h_adjoint_forward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "m70h_adjoint_reverse": (ctypes.c_int, [_vp, _vp, _vp, _dp]),
    "m70h_adjoint_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                                          ctypes.POINTER(ctypes.c_int)]),
    "m70h_adjoint_release": (ctypes.c_int, [_vp]),
    "m70h_adjoint_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "m70h_adjoint_batch_destroy": (None, [_vp]),
    "m70h_adjoint_batch_last_error": (ctypes.c_char_p, [_vp]),
    "m70h_adjoint_batch_forward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "m70h_adjoint_batch_reverse": (ctypes.c_int, [_vp, _vp, _vp, _dp]),
}

STATE_FIELDS = ("Zs", "D", "Qx", "Qy", "U", "V", "U_c", "V_c", "dQx_dx", "dQy_dy")


class Fields(ctypes.Structure):
    """m70h_fields (include/m70h.h): ten device pointers in state order."""
    _fields_ = [(k, ctypes.c_void_p) for k in STATE_FIELDS]


class M70HError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"m70h error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """m70h_params (include/m70h.h)."""
    _fields_ = [("drag_model", ctypes.c_int), ("visc_model", ctypes.c_int),
                ("offshore_bc", ctypes.c_int), ("dt", ctypes.c_double), ("Dmin", ctypes.c_double),
                ("alpha_f", ctypes.c_double), ("cD0", ctypes.c_double), ("nu0", ctypes.c_double),
                ("M", ctypes.c_double), ("ka", ctypes.c_double), ("z0", ctypes.c_double),
                ("alpha_r", ctypes.c_double)]


def load(path: str | None = None):
    """Loads libm70h.so, raising if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # torch's wheel ships its own libamdhip64.so.7 with the same SONAME as ROCm's; the first
    # copy loaded serves the whole process. Load torch's first (when installed) so that a
    # process using both torch and this library runs on one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m morfo70_amd.build` "
                           "(the Hydro step has no CPU fallback)")
    L = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def check(L, h, rc: int):
    if rc != M70H_OK:
        msg = L.m70h_last_error(h)
        raise M70HError(rc, msg.decode() if msg else "")
    return rc
