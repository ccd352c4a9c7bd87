"""ctypes binding of libmantaray_hip.so (include/mantaray_capi.h).

The library is the product: if it is missing this module raises, it never
falls back to another implementation.  torch, when importable, is imported
first so that the library binds to the same HIP runtime as torch (both name
it libamdhip64.so.7); device pointers then pass freely between the two.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmantaray_hip.so")

MR_OK = 0
MR_ERR_INVALID_ARGUMENT = -1
MR_ERR_IO = -2
MR_ERR_FORMAT = -3
MR_ERR_DEVICE = -4
MR_ERR_OUT_OF_MEMORY = -5
MR_ERR_STEP_COUNT = -6
MR_ERR_UNSUPPORTED = -7

MR_BATHY_CONSTANT, MR_BATHY_SLOPE, MR_BATHY_ARRAY, MR_BATHY_GRID = 0, 1, 2, 3
MR_CURRENT_CONSTANT, MR_CURRENT_GRID = 16, 17
MR_PREC_REFERENCE, MR_PREC_FP64_INTERP = 0, 1
MR_MATH_FAST, MR_MATH_LITERAL = 0, 1
MR_SOLOUT_STATE, MR_SOLOUT_NEXT_DERIV = 0, 1
MR_MEM_HOST, MR_MEM_DEVICE = 0, 1
MR_ADJOINT_SCRATCH_BYTES = 32 << 20

# every symbol include/mantaray_capi.h declares
EXPORTS = (
    "mr_options_init", "mr_last_error", "mr_version",
    "mr_bathy_open_nc3", "mr_current_open_nc3", "mr_bathy_from_arrays",
    "mr_current_from_arrays", "mr_bathy_constant", "mr_bathy_slope", "mr_bathy_array",
    "mr_current_constant", "mr_field_free", "mr_field_kind", "mr_field_grid_shape",
    "mr_field_copy_grid", "mr_field_upload", "mr_num_steps", "mr_trace_many",
    "mr_single_ray", "mr_rhs_many", "mr_checkpoint_count", "mr_adjoint_workspace_size",
    "mr_trace_fp64", "mr_adjoint_many", "mr_adjoint_many_noalloc", "mr_ray_result_json",
    "mr_ray_result_append_tsv",
    "mr_batch_create", "mr_batch_run", "mr_batch_info", "mr_batch_free",
)


class Scene(ctypes.Structure):
    _fields_ = [
        ("bathy", ctypes.c_void_p), ("current", ctypes.c_void_p),
        ("n_rays", ctypes.c_int64),
        ("x0", ctypes.c_void_p), ("y0", ctypes.c_void_p),
        ("kx0", ctypes.c_void_p), ("ky0", ctypes.c_void_p),
        ("t0", ctypes.c_double), ("t_end", ctypes.c_double), ("dt", ctypes.c_double),
        ("len_out", ctypes.c_void_p), ("end_out", ctypes.c_void_p),
    ]


class MantarayError(RuntimeError):
    """A nonzero status from the C ABI, with mr_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("math", ctypes.c_int32),
        ("solout", ctypes.c_int32),
        ("memory", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("record_tile", ctypes.c_int32),
    ]


_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_fieldp = ctypes.c_void_p
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_f32 = ctypes.c_float
_cs = ctypes.c_char_p

_SIGS = {
    "mr_options_init": (None, [ctypes.POINTER(Options)]),
    "mr_last_error": (ctypes.c_char_p, []),
    "mr_version": (ctypes.c_int, [_i32p, _i32p, _i32p]),
    "mr_bathy_open_nc3": (ctypes.c_int, [_cs, _cs, _cs, _cs, ctypes.POINTER(_fieldp)]),
    "mr_current_open_nc3": (ctypes.c_int, [_cs, _cs, _cs, _cs, _cs, ctypes.POINTER(_fieldp)]),
    "mr_bathy_from_arrays": (ctypes.c_int, [_i64, _i64, _vp, _vp, _i64, _vp,
                                            ctypes.POINTER(_fieldp)]),
    "mr_current_from_arrays": (ctypes.c_int, [_i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp,
                                              ctypes.POINTER(_fieldp)]),
    "mr_bathy_constant": (ctypes.c_int, [_f32, ctypes.POINTER(_fieldp)]),
    "mr_bathy_slope": (ctypes.c_int, [_f32, _f32, _f32, _f32, _f32, ctypes.POINTER(_fieldp)]),
    "mr_bathy_array": (ctypes.c_int, [_i64, _vp, ctypes.POINTER(_fieldp)]),
    "mr_current_constant": (ctypes.c_int, [_f64, _f64, ctypes.POINTER(_fieldp)]),
    "mr_field_free": (ctypes.c_int, [_fieldp]),
    "mr_field_kind": (ctypes.c_int, [_fieldp, _i32p]),
    "mr_field_grid_shape": (ctypes.c_int, [_fieldp, _i64p, _i64p, _i64p, _i64p]),
    "mr_field_copy_grid": (ctypes.c_int, [_fieldp, _vp, _vp, _vp, _vp]),
    "mr_field_upload": (ctypes.c_int, [_fieldp, ctypes.c_int32]),
    "mr_num_steps": (ctypes.c_int, [_f64, _f64, _f64, _i64p]),
    "mr_trace_many": (ctypes.c_int, [_fieldp, _fieldp, _i64, _vp, _vp, _vp, _vp, _f64, _f64,
                                     _f64, ctypes.POINTER(Options), _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp]),
    "mr_single_ray": (ctypes.c_int, [_fieldp, _fieldp, _f64, _f64, _f64, _f64, _f64, _f64,
                                     _f64, ctypes.POINTER(Options), _vp, _i64p]),
    "mr_rhs_many": (ctypes.c_int, [_fieldp, _fieldp, _i64, _vp, _vp, ctypes.POINTER(Options)]),
    "mr_checkpoint_count": (ctypes.c_int, [_f64, _f64, _f64, ctypes.c_int32, _i64p]),
    "mr_adjoint_workspace_size": (ctypes.c_int, [_i64, ctypes.c_int32, _i64p]),
    "mr_trace_fp64": (ctypes.c_int, [_fieldp, _fieldp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                                     _f64, _f64, _f64, ctypes.c_int32, ctypes.POINTER(Options),
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mr_adjoint_many": (ctypes.c_int, [_fieldp, _fieldp, _vp, _vp, _vp, _i64, _f64, _f64, _f64,
                                       ctypes.c_int32, ctypes.POINTER(Options), _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp]),
    "mr_adjoint_many_noalloc": (ctypes.c_int, [_fieldp, _fieldp, _vp, _vp, _vp, _i64, _f64, _f64,
                                               _f64, ctypes.c_int32, ctypes.POINTER(Options),
                                               _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp, _vp, _vp, _vp]),
    "mr_ray_result_json": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _i64,
                                          _i64p]),
    "mr_ray_result_append_tsv": (ctypes.c_int, [_cs, _vp, _vp, _vp, _vp, _vp, _i64, _i64]),
    "mr_batch_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(Scene),
                                       ctypes.POINTER(Options), ctypes.POINTER(_vp)]),
    "mr_batch_run": (ctypes.c_int, [_vp, ctypes.POINTER(Options)]),
    "mr_batch_info": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "mr_batch_free": (ctypes.c_int, [_vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libmantaray_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m mantaray_amd.build` "
            "(the HIP extension is required; there is no fallback)")
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != MR_OK:
        msg = lib().mr_last_error()
        raise MantarayError(rc, msg.decode() if msg else "")


def options(**kw) -> Options:
    o = Options()
    lib().mr_options_init(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
