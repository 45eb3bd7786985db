"""ctypes binding of libmfs_hip.so (include/mfs.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing, importing the binding raises, and every
solver object is a thin handle on the library's own C ABI.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfs_hip.so")

MFS_OK = 0
MFS_FP64 = 0
MFS_MIXED = 1

F_RHO, F_U, F_V, F_W, F_P, F_RES, F_DT, F_DQ, F_DQDX, F_DQDY, F_DQDZ, F_DIAG, F_QNODE = range(13)
FIELD_IDS = {
    "rho": F_RHO, "u": F_U, "v": F_V, "w": F_W, "p": F_P, "res": F_RES, "dt": F_DT,
    "dq": F_DQ, "dqdx": F_DQDX, "dqdy": F_DQDY, "dqdz": F_DQDZ, "diag": F_DIAG, "qnode": F_QNODE,
}


class MfsParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mach", "alpha", "beita", "re", "T", "gam", "gascon",
                                          "cfl_start", "cfl_end")] + [
        ("cfl_nstep", C.c_int)] + [(n, C.c_double) for n in (
            "epsa_r", "lhs_omga", "ratio_dtmax", "ratio_rhop_max", "ratio_rhop_min", "ratio_p_break", "amu",
            "ainf", "uqq", "u_s", "v_s", "w_s", "p_s", "rho_s",
            "p_bar", "gamm1", "p_stag", "e_stag", "p_min", "p_max", "rho_min", "rho_max", "p_break",
            "e_stag_max", "alf_l", "alf_n")]


class MfsGridInfo(C.Structure):
    _fields_ = [("nTCell", C.c_int), ("nTFace", C.c_int), ("nBFace", C.c_int), ("nTNode", C.c_int),
                ("preprocessed", C.c_int), ("decoupled", C.c_int)]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class MfsMeshView(C.Structure):
    _fields_ = [("nTCell", C.c_int), ("nTFace", C.c_int), ("nBFace", C.c_int), ("nTNode", C.c_int),
                ("x", _DP), ("y", _DP), ("z", _DP), ("nNPF", _IP), ("f2n", _IP), ("f2c", _IP),
                ("bc_type", _IP)] + [(n, _DP) for n in (
                    "xfn", "yfn", "zfn", "area", "xfc", "yfc", "zfc", "xcc", "ycc", "zcc", "vol",
                    "rho", "u", "v", "w", "p")]


class MfsScalars(C.Structure):
    _fields_ = [("iter_done", C.c_int), ("dt_max", C.c_double), ("dt_min", C.c_double),
                ("cfl_min", C.c_double)]


class MfsCtxInfo(C.Structure):
    _fields_ = [("nTCell", C.c_int), ("nTFace", C.c_int), ("nBFace", C.c_int), ("nTNode", C.c_int),
                ("precision", C.c_int), ("sweep_levels", C.c_int), ("device_bytes", C.c_size_t),
                ("tile_cells", C.c_int), ("sweep_waves", C.c_int), ("sweep_tiles", C.c_int),
                ("sweep_workgroups", C.c_int), ("sweep_by_tiles", C.c_int), ("compute_units", C.c_int),
                ("sweep_tail_chunks", C.c_int), ("sweep_tail_cells", C.c_int), ("sweep_tail_workgroups", C.c_int)]


class MfsCtxOptions(C.Structure):
    _fields_ = [("tile_cells", C.c_int), ("sweep_waves", C.c_int), ("sweep_by_tiles", C.c_int),
                ("gradient_stream", C.c_int), ("sweep_tail_grid", C.c_int)]


class MfsBatchOptions(C.Structure):
    _fields_ = [("gradient_stream", C.c_int)]


# Every symbol declared in include/mfs.h, with its ctypes signature.
_SIGS = {
    "mfs_params_default": (None, [C.POINTER(MfsParams)]),
    "mfs_params_derive": (C.c_int, [C.POINTER(MfsParams)]),
    "mfs_last_error": (C.c_char_p, []),
    "mfs_grid_generate_box_bump": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                             C.c_uint64, C.POINTER(C.c_void_p)]),
    "mfs_grid_generate_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, C.POINTER(C.c_void_p)]),
    "mfs_grid_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "mfs_grid_save": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "mfs_grid_decouple": (C.c_int, [C.c_void_p]),
    "mfs_grid_decouple_ex": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mfs_grid_preprocess_device": (C.c_int, [C.c_void_p, C.POINTER(MfsParams)]),
    "mfs_grid_preprocess": (C.c_int, [C.c_void_p, C.POINTER(MfsParams)]),
    "mfs_grid_get_info": (C.c_int, [C.c_void_p, C.POINTER(MfsGridInfo)]),
    "mfs_grid_array_bytes": (C.c_size_t, [C.c_void_p, C.c_char_p]),
    "mfs_grid_get_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]),
    "mfs_grid_destroy": (None, [C.c_void_p]),
    "mfs_grid_view": (C.c_int, [C.c_void_p, C.POINTER(MfsMeshView)]),
    "mfs_ctx_create": (C.c_int, [C.POINTER(MfsMeshView), C.POINTER(MfsParams), C.c_int, C.POINTER(C.c_void_p)]),
    "mfs_ctx_create_ex": (C.c_int, [C.POINTER(MfsMeshView), C.POINTER(MfsParams), C.c_int,
                                    C.POINTER(MfsCtxOptions), C.POINTER(C.c_void_p)]),
    "mfs_ctx_destroy": (C.c_int, [C.c_void_p]),
    "mfs_zero_residuals": (C.c_int, [C.c_void_p]),
    "mfs_inviscid_flux": (C.c_int, [C.c_void_p]),
    "mfs_compute_timestep": (C.c_int, [C.c_void_p]),
    "mfs_forward_lusgs": (C.c_int, [C.c_void_p]),
    "mfs_pass_interface_data": (C.c_int, [C.c_void_p]),
    "mfs_set_ghost_variables": (C.c_int, [C.c_void_p]),
    "mfs_ctx_set_ghost_refresh": (C.c_int, [C.c_void_p, C.c_int]),
    "mfs_calculate_gradient": (C.c_int, [C.c_void_p]),
    "mfs_iteration_done": (C.c_int, [C.c_void_p]),
    "mfs_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mfs_step_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "mfs_sync": (C.c_int, [C.c_void_p]),
    "mfs_history_count": (C.c_int, [C.c_void_p]),
    "mfs_get_history": (C.c_int, [C.c_void_p, _DP, C.c_int]),
    "mfs_field_count": (C.c_size_t, [C.c_void_p, C.c_int]),
    "mfs_download": (C.c_int, [C.c_void_p, C.c_int, _DP, C.c_size_t]),
    "mfs_upload": (C.c_int, [C.c_void_p, C.c_int, _DP, C.c_size_t]),
    "mfs_get_scalars": (C.c_int, [C.c_void_p, C.POINTER(MfsScalars)]),
    "mfs_ctx_get_info": (C.c_int, [C.c_void_p, C.POINTER(MfsCtxInfo)]),
    "mfs_ctx_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mfs_ctx_get_timing": (C.c_int, [C.c_void_p, C.c_char_p, _DP, _IP]),
    "mfs_ctx_reset_timing": (C.c_int, [C.c_void_p]),
    "mfs_batch_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mfs_batch_create_ex": (C.c_int, [C.c_int, C.POINTER(MfsBatchOptions), C.POINTER(C.c_void_p)]),
    "mfs_batch_destroy": (C.c_int, [C.c_void_p]),
    "mfs_batch_add": (C.c_int, [C.c_void_p, C.POINTER(MfsMeshView), C.POINTER(MfsParams), C.POINTER(C.c_int)]),
    "mfs_batch_attach": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "mfs_batch_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "mfs_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mfs_batch_sync": (C.c_int, [C.c_void_p]),
    "mfs_batch_ctx": (C.c_void_p, [C.c_void_p, C.c_int]),
    "mfs_batch_count": (C.c_int, [C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libmfs_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class MfsError(RuntimeError):
    def __init__(self, where: str, code: int):
        msg = lib().mfs_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def check(code: int, where: str) -> int:
    if code < 0:
        raise MfsError(where, code)
    return code
