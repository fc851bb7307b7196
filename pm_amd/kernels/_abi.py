"""ctypes mirror of the C ABI declared in include/pstm_hip.h.

The reference marshals kernel parameters through ctypes structures mirrored
by hand in Metal shader structs (pstm/kernels/metal_compiled.py:137-289); its
CurvedRayParams and VTI bindings drifted out of sync with the shaders.  Here
the layout is declared once in C, mirrored once below, and checked against the
library's own ``sizeof`` at load time (``check_layout``).
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

ABI_VERSION = 5

TT_STRAIGHT, TT_CURVED, TT_VTI, TT_VTI_LEARNED, TT_CURVED_LEARNED = 0, 1, 2, 3, 4
AA_NONE, AA_TRIANGLE = 0, 1
VEL_1D, VEL_3D = 0, 1
ETA_CONST, ETA_1D, ETA_3D = 0, 1, 2
#: interpolation codes; names and aliases as the reference's get_method_code
#: (pstm/kernels/interpolation.py:322-356), unknown names -> linear
INTERP_LINEAR, INTERP_NEAREST, INTERP_CUBIC = 0, 1, 2
INTERP_SINC4, INTERP_SINC8, INTERP_SINC16, INTERP_LANCZOS3, INTERP_LANCZOS5 = 3, 4, 5, 6, 7
INTERP_CODES = {"nearest": INTERP_NEAREST, "linear": INTERP_LINEAR, "cubic": INTERP_CUBIC, "spline": INTERP_CUBIC,
                "sinc4": INTERP_SINC4, "sinc8": INTERP_SINC8, "sinc": INTERP_SINC8, "sinc16": INTERP_SINC16,
                "lanczos3": INTERP_LANCZOS3, "lanczos": INTERP_LANCZOS3, "lanczos5": INTERP_LANCZOS5}


def interp_code(method) -> int:
    """Code of an interpolation method name (or enum with a .value), linear
    for None or an unknown name, as the reference maps them."""
    if method is None:
        return INTERP_LINEAR
    name = getattr(method, "value", method)
    return INTERP_CODES.get(str(name).lower().strip(), INTERP_LINEAR)


SEL_OFFSET_MIN, SEL_OFFSET_MAX = 1, 2
SEL_OVT_X_MIN, SEL_OVT_X_MAX, SEL_OVT_Y_MIN, SEL_OVT_Y_MAX = 4, 8, 16, 32

_vp = ctypes.c_void_p


class PstmMigrateArgs(ctypes.Structure):
    """Field-for-field mirror of ``struct PstmMigrateArgs``."""

    _fields_ = [
        ("amplitudes", _vp), ("source_x", _vp), ("source_y", _vp),
        ("receiver_x", _vp), ("receiver_y", _vp),
        ("midpoint_x", _vp), ("midpoint_y", _vp),
        ("trace_weights", _vp), ("trace_index", _vp),
        ("x_axis", _vp), ("y_axis", _vp), ("t_axis_ms", _vp),
        ("vrms", _vp), ("eta", _vp),
        ("image", _vp), ("fold", _vp), ("n_effective", _vp),
        ("n_traces", ctypes.c_int32), ("n_samples", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("vel_mode", ctypes.c_int32), ("eta_mode", ctypes.c_int32),
        ("traveltime", ctypes.c_int32), ("aa_mode", ctypes.c_int32),
        ("apply_spreading", ctypes.c_int32),
        ("apply_obliquity", ctypes.c_int32),
        ("aa_max_half_width", ctypes.c_int32),
        ("dt_in_ms", ctypes.c_double), ("t_start_in_ms", ctypes.c_double),
        ("max_aperture_m", ctypes.c_double),
        ("min_aperture_m", ctypes.c_double),
        ("max_dip_deg", ctypes.c_double), ("taper_fraction", ctypes.c_double),
        ("grid_dx", ctypes.c_double), ("grid_dy", ctypes.c_double),
        ("curved_v0", ctypes.c_double), ("curved_k", ctypes.c_double),
        ("eta_constant", ctypes.c_double),
        ("mlp_params", _vp), ("mlp_node_spacing", ctypes.c_int32), ("interpolation", ctypes.c_int32),
        ("mlp_tau_max", ctypes.c_double), ("mlp_v_lo", ctypes.c_double), ("mlp_v_hi", ctypes.c_double),
        ("mlp_eta_max", ctypes.c_double), ("mlp_out_scale", ctypes.c_double),
        ("mlp_weight_terms", ctypes.c_int32), ("mlp_node_growth", ctypes.c_int32),
        ("mlp_node_max", ctypes.c_int32), ("reserved0", ctypes.c_int32),
    ]


class PstmTileDesc(ctypes.Structure):
    """Field-for-field mirror of ``struct PstmTileDesc``."""

    _fields_ = [
        ("x_axis", _vp), ("y_axis", _vp), ("vrms", _vp), ("eta", _vp),
        ("image", _vp), ("fold", _vp), ("trace_index", _vp),
        ("n_traces", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


#: every symbol include/pstm_hip.h declares
EXPORTED_SYMBOLS = (
    "pstm_hip_abi_version", "pstm_hip_args_size", "pstm_hip_last_error",
    "pstm_hip_init", "pstm_hip_release", "pstm_hip_stream_sync",
    "pstm_hip_migrate_tile", "pstm_hip_migrate_tile_noalloc", "pstm_hip_fold_tile", "pstm_hip_migrate_bins",
    "pstm_hip_profile_reset", "pstm_hip_profile_read",
    "pstm_hip_set_tuning", "pstm_hip_set_pipeline", "pstm_hip_last_launch",
    "pstm_hip_plan_create", "pstm_hip_plan_run", "pstm_hip_plan_info", "pstm_hip_plan_destroy",
    "pstm_hip_bin_headers", "pstm_hip_select_rows", "pstm_hip_fold_normalize", "pstm_hip_resample_uniform",
    "pstm_mlp_param_count", "pstm_mlp_last_error", "pstm_mlp_set_precision", "pstm_mlp_forward",
    "pstm_mlp_forward_mfma", "pstm_mlp_forward_mfma_terms", "pstm_mlp_loss_grad",
    "pstm_mlp_adam", "pstm_mlp_train_step", "pstm_mlp_train", "pstm_mlp_loss_grad_fused", "pstm_vti_leg_times",
)

MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 5, 64, 2, 12994

PACKAGE_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PACKAGE_DIR / "lib" / "libpstm_hip.so"


def library_path() -> Path:
    return Path(os.environ.get("PSTM_HIP_LIB", str(LIB_PATH)))


def declare(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Attach argument/return types to every exported entry point."""
    P = ctypes.POINTER(PstmMigrateArgs)
    lib.pstm_hip_abi_version.restype = ctypes.c_int
    lib.pstm_hip_abi_version.argtypes = []
    lib.pstm_hip_args_size.restype = ctypes.c_size_t
    lib.pstm_hip_args_size.argtypes = []
    lib.pstm_hip_last_error.restype = ctypes.c_char_p
    lib.pstm_hip_last_error.argtypes = []
    lib.pstm_hip_init.restype = ctypes.c_int
    lib.pstm_hip_init.argtypes = [ctypes.c_int]
    lib.pstm_hip_release.restype = ctypes.c_int
    lib.pstm_hip_release.argtypes = []
    lib.pstm_hip_stream_sync.restype = ctypes.c_int
    lib.pstm_hip_stream_sync.argtypes = [_vp]
    lib.pstm_hip_migrate_tile.restype = ctypes.c_int
    lib.pstm_hip_migrate_tile.argtypes = [P, _vp]
    lib.pstm_hip_migrate_tile_noalloc.restype = ctypes.c_int
    lib.pstm_hip_migrate_tile_noalloc.argtypes = [P, _vp]
    lib.pstm_hip_fold_tile.restype = ctypes.c_int
    lib.pstm_hip_fold_tile.argtypes = [P, _vp]
    lib.pstm_hip_migrate_bins.restype = ctypes.c_int
    lib.pstm_hip_migrate_bins.argtypes = [P, _vp, ctypes.c_int32, _vp]
    lib.pstm_hip_profile_reset.restype = ctypes.c_int
    lib.pstm_hip_profile_reset.argtypes = [ctypes.c_int]
    lib.pstm_hip_profile_read.restype = ctypes.c_int
    lib.pstm_hip_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_longlong)]
    lib.pstm_hip_set_tuning.restype = ctypes.c_int
    lib.pstm_hip_set_tuning.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.pstm_hip_set_pipeline.restype = ctypes.c_int
    lib.pstm_hip_set_pipeline.argtypes = [ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.pstm_hip_last_launch.restype = ctypes.c_int
    lib.pstm_hip_last_launch.argtypes = [ip, ip, ip, ip]
    lib.pstm_hip_plan_create.restype = ctypes.c_int
    lib.pstm_hip_plan_create.argtypes = [_vp, _vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    lib.pstm_hip_plan_run.restype = ctypes.c_int
    lib.pstm_hip_plan_run.argtypes = [_vp, ctypes.c_int32, _vp]
    lib.pstm_hip_plan_info.restype = ctypes.c_int
    lib.pstm_hip_plan_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong), ip, ip, ip]
    lib.pstm_hip_plan_destroy.restype = ctypes.c_int
    lib.pstm_hip_plan_destroy.argtypes = [_vp]
    d, i32 = ctypes.c_double, ctypes.c_int32
    lib.pstm_hip_bin_headers.restype = ctypes.c_int
    lib.pstm_hip_bin_headers.argtypes = [_vp, _vp, _vp, _vp, _vp, i32, d, d, d, d, i32, i32, i32, i32, i32, i32,
                                         _vp, _vp, _vp, _vp, _vp]
    lib.pstm_hip_select_rows.restype = ctypes.c_int
    lib.pstm_hip_select_rows.argtypes = [_vp, _vp, _vp, _vp, _vp, i32, i32, d, d, d, d, d, d, _vp, _vp]
    lib.pstm_hip_fold_normalize.restype = ctypes.c_int
    lib.pstm_hip_fold_normalize.argtypes = [_vp, _vp, ctypes.c_longlong, i32, _vp, _vp]
    lib.pstm_hip_resample_uniform.restype = ctypes.c_int
    lib.pstm_hip_resample_uniform.argtypes = [_vp, _vp, i32, _vp, i32, ctypes.c_longlong, _vp, _vp]
    f = ctypes.c_float
    lib.pstm_mlp_param_count.restype = ctypes.c_int
    lib.pstm_mlp_param_count.argtypes = []
    lib.pstm_mlp_last_error.restype = ctypes.c_char_p
    lib.pstm_mlp_last_error.argtypes = []
    lib.pstm_mlp_set_precision.restype = ctypes.c_int
    lib.pstm_mlp_set_precision.argtypes = [ctypes.c_int]
    lib.pstm_mlp_forward.restype = ctypes.c_int
    lib.pstm_mlp_forward.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
    lib.pstm_mlp_forward_mfma.restype = ctypes.c_int
    lib.pstm_mlp_forward_mfma.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
    lib.pstm_mlp_forward_mfma_terms.restype = ctypes.c_int
    lib.pstm_mlp_forward_mfma_terms.argtypes = [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]
    lib.pstm_mlp_loss_grad.restype = ctypes.c_int
    lib.pstm_mlp_loss_grad.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
    lib.pstm_mlp_adam.restype = ctypes.c_int
    lib.pstm_mlp_adam.argtypes = [_vp, _vp, _vp, _vp, f, f, f, f, ctypes.c_int, _vp]
    lib.pstm_mlp_train_step.restype = ctypes.c_int
    lib.pstm_mlp_train_step.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, f, f, f, f,
                                        ctypes.c_int, _vp, _vp]
    lib.pstm_mlp_train.restype = ctypes.c_int
    lib.pstm_mlp_train.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, f, f,
                                   f, f, f, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, _vp, _vp]
    lib.pstm_mlp_loss_grad_fused.restype = ctypes.c_int
    lib.pstm_mlp_loss_grad_fused.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
    lib.pstm_vti_leg_times.restype = ctypes.c_int
    lib.pstm_vti_leg_times.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_longlong, _vp, _vp]
    return lib


def last_launch(lib: ctypes.CDLL) -> dict:
    vals = [ctypes.c_int() for _ in range(4)]
    lib.pstm_hip_last_launch(*[ctypes.byref(v) for v in vals])
    return dict(zip(("mw", "batch", "grid", "items"), (v.value for v in vals)))


def check_layout(lib: ctypes.CDLL) -> None:
    """Refuse a library whose ABI differs from this mirror."""
    ver = lib.pstm_hip_abi_version()
    if ver != ABI_VERSION:
        raise RuntimeError(f"libpstm_hip ABI version {ver} != {ABI_VERSION}")
    if lib.pstm_mlp_param_count() != MLP_PARAMS:
        raise RuntimeError("libpstm_hip MLP parameter count mismatch")
    size = lib.pstm_hip_args_size()
    if size != ctypes.sizeof(PstmMigrateArgs):
        raise RuntimeError(
            f"PstmMigrateArgs layout mismatch: C sizeof {size} != ctypes "
            f"{ctypes.sizeof(PstmMigrateArgs)}")


_LIB: ctypes.CDLL | None = None


def load_library() -> ctypes.CDLL:
    """Load the built HIP library (no device work happens at load time).

    Raises RuntimeError when the library has not been built: the backend has no
    CPU fallback.
    """
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise RuntimeError(
            f"HIP backend library not built: {path} is missing "
            "(build it with python -m pm_amd.build)")
    lib = declare(ctypes.CDLL(str(path)))
    check_layout(lib)
    _LIB = lib
    return lib


def last_error(lib: ctypes.CDLL) -> str:
    msg = lib.pstm_hip_last_error()
    return msg.decode() if msg else ""
