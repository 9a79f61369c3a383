"""ctypes declarations of the C ABI in include/mif_gpu.h.

The library is loaded from the package's in-tree _build/ directory; there is
no fallback: if it is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libmif_gpu.so")

# mif_status
OK = 0
STATUS_NAMES = {
    0: "MIF_OK", 1: "MIF_ERR_INVALID_ARGUMENT", 2: "MIF_ERR_HIP", 3: "MIF_ERR_OUT_OF_MEMORY",
    4: "MIF_ERR_NON_FINITE", 5: "MIF_ERR_UNSUPPORTED", 6: "MIF_ERR_NO_DEVICE",
}

# mif_staggering
STAGGER_X, STAGGER_Y, STAGGER_Z, STAGGER_NONE = 0, 1, 2, 3

# mif_precision
FP64, MIXED = 0, 1

# mif_solve_precision
SOLVE_AUTO, SOLVE_FP32, SOLVE_FP64 = 0, 1, 2

# mif_defect_schedule
DEFECT_EVERY_STAGE, DEFECT_LAST_STAGE = 0, 1

# mif_fn_kind
FN_ZERO = 0
FN_ES_VELOCITY = 1
FN_ES_PRESSURE = 2
FN_ES_GRAD_P = 3
FN_VT_VELOCITY = 4
FN_VT_FORCING = 5
FN_PT_VELOCITY = 6
FN_PT_PRESSURE = 7
FN_PT_GRAD_P = 8
FN_LID_1 = 9
FN_LID_2 = 10
FN_CONSTANT = 11
FN_TIME_DIFFERENCE = 12

# mif_kernel_class
KERNEL_CLASSES = ["stage", "bc", "div_rhs", "fft_x", "fft_y", "fft_z_fused", "unpack", "project", "defect",
                  "fft_x_div", "fft_x_unpack", "adjoint", "p_rim"]


class MifGrid(ctypes.Structure):
    _fields_ = [
        ("Nx_global", ctypes.c_uint64), ("Ny_global", ctypes.c_uint64), ("Nz_global", ctypes.c_uint64),
        ("x_size", ctypes.c_double), ("y_size", ctypes.c_double), ("z_size", ctypes.c_double),
        ("min_x", ctypes.c_double), ("min_y", ctypes.c_double), ("min_z", ctypes.c_double),
        ("Re", ctypes.c_double), ("final_time", ctypes.c_double), ("num_time_steps", ctypes.c_uint32),
        ("periodic", ctypes.c_int32 * 3),
    ]


class MifFunction(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("base", ctypes.c_int32), ("Re", ctypes.c_double),
        ("c", ctypes.c_double * 3), ("t0", ctypes.c_double), ("t1", ctypes.c_double),
    ]


# every exported symbol with its argument types (restype int unless noted)
_VP = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_DP = ctypes.POINTER(ctypes.c_double)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_FP = ctypes.POINTER(MifFunction)
_V3 = ctypes.c_void_p * 3

SIGNATURES = {
    "mif_ctx_create": [ctypes.POINTER(MifGrid), _I, _I, ctypes.POINTER(_VP)],
    "mif_ctx_create_cell_centred": [ctypes.POINTER(MifGrid), ctypes.POINTER(_I), _I, _I, ctypes.POINTER(_VP)],
    "mif_ctx_destroy": [_VP],
    "mif_ctx_synchronize": [_VP],
    "mif_last_error": [_VP],
    "mif_ctx_info": [_VP, _DP, _U64P],
    "mif_ctx_set_defect_sweeps": [_VP, _I],
    "mif_ctx_get_defect_sweeps": [_VP, ctypes.POINTER(_I)],
    "mif_ctx_set_defect_schedule": [_VP, _I],
    "mif_ctx_set_solve_precision": [_VP, _I],
    "mif_ctx_get_solve_precision": [_VP, ctypes.POINTER(_I), _DP],
    "mif_field_create": [_VP, _I, ctypes.POINTER(_VP)],
    "mif_field_destroy": [_VP],
    "mif_field_extents": [_VP, _U64P],
    "mif_field_upload": [_VP, _DP],
    "mif_field_download": [_VP, _DP],
    "mif_field_upload_device": [_VP, _VP],
    "mif_field_download_device": [_VP, _VP],
    "mif_field_swap": [_VP, _VP],
    "mif_field_set": [_VP, _FP, _I, _D, _I],
    "mif_field_apply_periodic_bc": [_VP],
    "mif_velocity_apply_bc": [_V3, _FP, _D],
    "mif_timestep": [_V3, _V3, _V3, _FP, _D, _VP, _VP],
    "mif_timestep_n": [_V3, _V3, _V3, _FP, _D, _VP, _VP, _I],
    "mif_timestep_nhn": [_V3, _V3, _V3, _FP, _FP, _D, _VP, _VP],
    "mif_timestep_velocity": [_V3, _V3, _V3, _FP, _FP, _D],
    "mif_solve_pressure_homogeneous_periodic": [_VP, _V3, _D],
    "mif_solve_pressure_non_homogeneous_neumann": [_VP, _V3, _FP, _D, _D],
    "mif_poisson_solve": [_VP],
    "mif_poisson_set_rhs": [_VP, _DP],
    "mif_poisson_run": [_VP],
    "mif_poisson_get_solution": [_VP, _DP],
    "mif_adjust_pressure": [_VP, _FP, _D],
    "mif_adjust_pressure_tabulated": [_VP, _VP],
    "mif_error_norms_velocity": [_V3, _FP, _D, _DP],
    "mif_error_norms_velocity_tabulated": [_V3, _V3, _DP],
    "mif_error_norms_pressure": [_VP, _FP, _D, _DP],
    "mif_error_norms_pressure_tabulated": [_VP, _VP, _DP],
    "mif_pressure_gradient": [_VP, _V3],
    "mif_ctx_set_body_force": [_VP, _DP],
    "mif_ctx_set_stage_records": [_VP, _I],
    "mif_timestep_adjoint": [_V3, _V3, _V3, _V3, _VP, _DP],
    "mif_adjoint_force_take": [_VP, _DP],
    "mif_field_copy": [_VP, _VP],
    "mif_field_dot": [_VP, _VP, _DP],
    "mif_field_gather": [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, _DP],
    "mif_field_gather_noalloc": [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, _DP],
    "mif_bc_table_size": [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "mif_velocity_apply_bc_tabulated": [_V3, _DP],
    "mif_timestep_tabulated": [_V3, _V3, _V3, _DP, _D, _VP, _VP],
    "mif_timestep_nhn_tabulated": [_V3, _V3, _V3, _DP, _DP, _D, _VP, _VP],
    "mif_timestep_velocity_tabulated": [_V3, _V3, _V3, _DP, _FP, _D],
    "mif_solve_pressure_non_homogeneous_neumann_tabulated": [_VP, _V3, _DP, _D],
    "mif_ctx_set_timing": [_VP, _I],
    "mif_ctx_kernel_time": [_VP, _I, _DP, _U64P],
    "mif_kernel_class_name": [_I],
    "mif_stream_probe": [_VP, _I, _I, ctypes.c_int64, _I, _DP],
}
RESTYPES = {"mif_last_error": ctypes.c_char_p, "mif_kernel_class_name": ctypes.c_char_p}

_lib = None


def load() -> ctypes.CDLL:
    """Load libmif_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        # One HIP runtime per process.  torch bundles its own libamdhip64
        # (SONAME libamdhip64.so.7) and reaches it by the unversioned name
        # through its RPATH.  When torch is loaded first, this library's
        # libamdhip64.so.7 dependency binds to that copy; loaded the other
        # way round the process would hold two runtimes, and torch's would
        # see no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = RESTYPES.get(name, ctypes.c_int)
        _lib = lib
    return _lib
