"""Loader and ctypes signatures of libmexct.so (the HIP operators' C ABI).

The library is built in-tree by mex-cuda-ct_amd/build.py.  There is no CPU
fallback: if the library is missing every operator raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MEXCT_LIB names another build of the library (the sanitizer build, for one)
LIB_PATH = os.environ.get("MEXCT_LIB") or os.path.join(_PKG, "lib", "libmexct.so")

OK, EINVAL, EGEOM, EHIP, ENOMEM, ENOTSUP = 0, -1, -2, -3, -4, -5
F32, F16, BF16 = 0, 1, 2


class MexctError(RuntimeError):
    """A negative return code of the C ABI with mexct_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


class Opts(C.Structure):
    _fields_ = [("storage", C.c_int32), ("tv_steps", C.c_int32), ("zrange", C.c_int32),
                ("long_ray", C.c_int32), ("g3d_bp", C.c_int32)]


G3D_BRICKS, G3D_DETERMINISTIC, G3D_ATOMICS = 0, 1, 2


class Request(C.Structure):
    """mexct_request: one request of the continuous-batching server."""
    _fields_ = [("op", C.c_int32), ("geom", C.c_int32), ("storage", C.c_int32),
                ("window", C.c_int32), ("W", C.c_int64), ("H", C.c_int64), ("nProj", C.c_int64),
                ("sad", C.c_float), ("osc_tv", C.c_float), ("cutoff", C.c_double),
                ("osc_p", C.c_double), ("osc_iters", C.c_int32), ("osc_s_init", C.c_int32),
                ("osc_s_final", C.c_int32), ("tv_steps", C.c_int32), ("angles", C.c_void_p),
                ("input", C.c_void_p), ("input2", C.c_void_p), ("output", C.c_void_p),
                ("first_picks", C.c_void_p), ("resident", C.c_int32)]


OP_FP, OP_BP, OP_FDK, OP_OSCTV = 0, 1, 2, 3

_vp, _i32, _i64, _f, _d = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
_op = C.POINTER(Opts)

# name -> argtypes (restype int unless noted)
SIGNATURES = {
    "mexct_version": [],
    "mexct_last_error": [],
    "mexct_ctx_create": [C.POINTER(_vp), C.c_int, _vp],
    "mexct_ctx_destroy": [_vp],
    "mexct_ctx_sync": [_vp],
    "mexct_ctx_workspace_bytes": [_vp],
    "mexct_fp": [_vp, _i64, _i64, _f, _vp, _i64, _i32, _i32, _vp, _op],
    "mexct_fp_dev": [_vp, _vp, _i32, _i64, _i64, _f, _vp, _i32, _i32, _vp, _op],
    "mexct_bp": [_vp, _i64, _i64, _i64, _f, _vp, _i64, _i32, _vp, _op],
    "mexct_bp_dev": [_vp, _vp, _i32, _i64, _i64, _i64, _f, _vp, _i32, _vp, _op],
    "mexct_fp_loads_dev": [_vp, _vp, _i32, _i64, _i64, _f, _vp, _i32, _i32, _vp, _op],
    "mexct_bp_loads_dev": [_vp, _vp, _i32, _i64, _i64, _i64, _f, _vp, _i32, _vp, _op],
    "mexct_fp_batch_dev": [_vp, _vp, _i32, _i32, _i64, _i64, _f, _vp, _i32, _i32, _vp, _op],
    "mexct_bp_batch_dev": [_vp, _vp, _i32, _i32, _i64, _i64, _i64, _f, _vp, _i32, _vp, _op],
    "mexct_set_autotune": [_i32],
    "mexct_ctx_set_autotune": [_vp, _i32],
    "mexct_tuned_config": [_i32, _i32, _i32, _i64, _i64, _i64, _vp],
    "mexct_g3d_fp": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _i32, _i32, _vp, _op],
    "mexct_g3d_fp_dev": [_vp, _vp, _i32, _i64, _i64, _vp, _vp, _i64, _i64, _i32, _i32, _vp, _op],
    "mexct_g3d_bp": [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i32, _i64, _i64, _vp, _op],
    "mexct_g3d_bp_dev": [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _i32, _i64, _i64, _vp, _op],
    "mexct_osc_iter": [_vp, _i64, _i64, _vp, _vp, _i64, _i32, _vp, _i64, _vp, _i64, _f, _i32, _op],
    "mexct_osc_iter_dev": [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _i32, _vp, _vp, _f, _i32, _op],
    "mexct_tvmin": [_vp, _i64, _i64, _i64, _f, _vp, _op],
    "mexct_tvmin_dev": [_vp, _vp, _i64, _i64, _i64, _f, _op],
    "mexct_prefilter": [_vp, _i64, _i64, _i64, _i32, _d, _vp, _i64, _i32, _d, _d, _vp],
    "mexct_prefilter_dev": [_vp, _vp, _i64, _i64, _i64, _i32, _d, _vp, _i32, _d, _d, _vp],
    "mexct_prefilter_noalloc_dev": [_vp, _vp, _i64, _i64, _i64, _i32, _d, _vp, _i32, _d, _d, _vp],
    "mexct_parker_weight": [_vp, _i64, _i64, _i64, _vp, _i64, _i32, _d, _d, _vp],
    "mexct_parker_weight_dev": [_vp, _vp, _i64, _i64, _i64, _vp, _i32, _d, _d, _vp],
    "mexct_fbp_recon": [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _i32, _d, _i32, _d, _d, _vp, _op],
    "mexct_fbp_recon_dev": [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _d, _i32, _d, _d, _vp,
                            _op],
    "mexct_osc_tv_recon": [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _f, _i32, _i32, _f, _i32, _i32,
                           _d, _vp, _vp, _op],
    "mexct_osc_tv_recon_dev": [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _f, _i32, _i32, _f, _i32,
                               _i32, _d, _vp, _vp, _op],
    "mexct_osc_tv_iterations": [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _f, _i32, _i32, _f, _i32,
                                _i32, _d, _vp, _i32, _i32, _vp, _op],
    "mexct_osc_tv_iterations_dev": [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _f, _i32, _i32, _f,
                                    _i32, _i32, _d, _vp, _i32, _i32, _vp, _op],
    "mexct_probe_gather_dev": [_vp, _i32, _i64, _i64, _i32, _i32, C.POINTER(_i64)],
    "mexct_server_create": [C.POINTER(_vp), C.c_int, _i64, _i32, _i32],
    "mexct_server_submit": [_vp, C.POINTER(Request), C.POINTER(C.c_uint64)],
    "mexct_server_footprint": [C.POINTER(Request)],
    "mexct_server_wait": [_vp, C.c_uint64, C.POINTER(_i32), C.POINTER(_d), C.POINTER(_d)],
    "mexct_server_stats": [_vp, C.POINTER(_i64), _i32],
    "mexct_server_destroy": [_vp],
    "mexct_calc_num_subsets": [_i32, _i32, _d, _i32, _i32],
    "mexct_calc_subset_order": [_i32, _i32, _vp],
}

_lib = None


def lib():
    """Load libmexct.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MexctError(ENOTSUP, f"{LIB_PATH} is missing: run mex-cuda-ct_amd/build.py "
                                      "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            # a symbol the built library lacks stays unset (its callers raise
            # AttributeError); the export test checks that none is missing
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = C.c_int
        L.mexct_last_error.restype = C.c_char_p
        L.mexct_ctx_workspace_bytes.restype = C.c_int64
        L.mexct_server_footprint.restype = C.c_int64
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise MexctError(rc, lib().mexct_last_error().decode(errors="replace"))


def opts(storage: int = F32, tv_steps: int = 0, zrange: int = 0, long_ray: int = 0,
         g3d_bp: int = G3D_BRICKS):
    return C.byref(Opts(storage, tv_steps, zrange, long_ray, g3d_bp))
