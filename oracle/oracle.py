"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker.  The product
package never imports this module.

Arrays follow MATLAB's column-major convention: a volume is a numpy array
of shape (W, W, H) in Fortran order (vol[x, y, z]), a sinogram has shape
(W, H, nProj).  Every function returns new arrays and leaves its inputs
untouched, except the *_inplace OSC helpers which mirror CUDAmex_oscIter.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MEXCT_ORACLE_LIB: another build of the oracle (the sanitizer build)
_LIB_PATH = os.environ.get("MEXCT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

CLAMP, BORDER = 0, 1
WINDOWS = {"ramp": 0, "shepp-logan": 1, "cosine": 2, "hamming": 3, "hann": 4, "none": 5}

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i, f, d, i64 = C.c_int, C.c_float, C.c_double, C.c_int64
        sig = {
            "orc_set_threads": (None, [i]),
            "orc_get_threads": (i, []),
            "orc_set_texture_weight_bits": (None, [i]),
            "orc_get_texture_weight_bits": (i, []),
            "orc_tex3": (f, [_f32p, i, i, i, f, f, f, i]),
            "orc_tex2": (f, [_f32p, i, i, f, f, i]),
            "orc_fp": (i64, [_f32p, i, i, f, _f32p, i, i, _f32p]),
            "orc_fp_f64": (i64, [_f32p, i, i, d, _f32p, i, i, _f64p]),
            "orc_fp_f64_rays": (None, [_f32p, i, i, d, _f32p, i, i, _i32p, _i32p, _i32p, _f64p]),
            "orc_fp_rays": (None, [_f32p, i, i, f, _f32p, i, i, _i32p, _i32p, _i32p, _f32p]),
            "orc_bp_cols": (None, [_f32p, i, i, i, f, _f32p, i, i, _i32p, _i32p, _f32p]),
            "orc_bp": (None, [_f32p, i, i, i, f, _f32p, i, _f32p]),
            "orc_bp_f64": (None, [_f32p, i, i, i, d, _f32p, i, _f64p]),
            "orc_bp_f64_cols": (None, [_f32p, i, i, i, d, _f32p, i, i, _i32p, _i32p, _f64p]),
            "orc_g3d_fp": (i64, [_f32p, i, i, _f32p, _f32p, i, i, i, i, _f32p]),
            "orc_g3d_fp_f64": (None, [_f32p, i, i, _f32p, _f32p, i, i, i, i, _f64p]),
            "orc_g3d_bp": (None, [_f32p, i, i, i, _f32p, _f32p, i, i, i, _f32p]),
            "orc_g3d_bp_f64": (None, [_f32p, i, i, i, _f32p, _f32p, i, i, i, _f64p]),
            "orc_osc_iter": (i, [_f32p, i, i, _f32p, _f32p, i, i, _i32p, _f32p, f, i]),
            "orc_osc_iter_f64": (i, [_f64p, i, i, _f64p, _f64p, i, i, _i32p, _f32p, d, i]),
            "orc_tvmin": (None, [_f32p, i, i, i, f, i]),
            "orc_tvmin_f64": (None, [_f64p, i, i, i, d, i]),
            "orc_tv_gradient": (None, [_f32p, i, i, i, _f32p]),
            "orc_design_filter": (i, [i, i, d, C.c_void_p]),
            "orc_prefilter": (i, [_f32p, i, i, i, i, d, _f32p, i, d, d, _f32p]),
            "orc_prefilter_f64": (i, [_f32p, i, i, i, i, d, _f32p, i, d, d, _f64p]),
            "orc_parker_window": (i, [i, i, i, d, _f32p, d, _f32p]),
            "orc_log_transform": (None, [_f32p, _f32p, i64, _f32p]),
            "orc_calc_num_subsets": (i, [i, i, d, i, i]),
            "orc_calc_subset_order": (None, [i, i, _i32p]),
            "orc_osc_tv_recon": (i, [_f32p, _f32p, i, i, i, _f32p, f, i, i, f, i, i, d, i, _i32p,
                                     _f32p]),
            "orc_osc_tv_recon_f64": (i, [_f32p, _f32p, i, i, i, _f32p, f, i, i, f, i, i, d, i,
                                         _i32p, _f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class use_library:
    """Context manager: every wrapper below calls the oracle built at `path`
    (same exports) instead of liboracle.so.  The tests use it to run a pin
    against a copy of the oracle with a deliberate mistake planted in it."""

    def __init__(self, path):
        self.path = path
        self.saved = None

    def __enter__(self):
        global _lib, _LIB_PATH
        self.saved = (_lib, _LIB_PATH)
        _lib, _LIB_PATH = None, self.path
        return self

    def __exit__(self, *exc):
        global _lib, _LIB_PATH
        _lib, _LIB_PATH = self.saved
        return False


def _f32(a):
    return np.asfortranarray(a, dtype=np.float32)


def _angles(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel()).reshape(-1, order="F")


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def threads() -> int:
    return lib().orc_get_threads()


class texture_weight_bits:
    """Context manager: every sampler call inside rounds its filter
    fractions to `bits` fraction bits (8: the texture unit's documented
    9-bit fixed-point weights; 0: exact FP32 weights, the default)."""

    def __init__(self, bits):
        self.bits = int(bits)
        self.saved = 0

    def __enter__(self):
        self.saved = lib().orc_get_texture_weight_bits()
        lib().orc_set_texture_weight_bits(self.bits)
        return self

    def __exit__(self, *exc):
        lib().orc_set_texture_weight_bits(self.saved)
        return False


def tex3(T, x, y, z, mode=CLAMP):
    T = _f32(T)
    return lib().orc_tex3(T, T.shape[0], T.shape[1], T.shape[2], x, y, z, mode)


def tex2(T, x, y, mode=CLAMP):
    T = _f32(T)
    return lib().orc_tex2(T, T.shape[0], T.shape[1], x, y, mode)


def fp(vol, sad, angles, nproj, geom, f64=False, return_samples=False):
    """CUDAmex_FP(vol, SAD, angles, nProj, geomFlag) -> [W, H, nProj]."""
    vol = _f32(vol)
    W, H = vol.shape[0], vol.shape[2]
    ang = _angles(angles)
    assert ang.size >= nproj
    if f64:
        out = np.zeros((W, H, nproj), dtype=np.float64, order="F")
        n = lib().orc_fp_f64(vol, W, H, float(sad), ang, nproj, geom, out)
    else:
        out = np.zeros((W, H, nproj), dtype=np.float32, order="F")
        n = lib().orc_fp(vol, W, H, float(sad), ang, nproj, geom, out)
    return (out, n) if return_samples else out


def fp_f64_rays(vol, sad, angles, geom, ix, iy, view):
    vol = _f32(vol)
    W, H = vol.shape[0], vol.shape[2]
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    iy = np.ascontiguousarray(iy, dtype=np.int32)
    view = np.ascontiguousarray(view, dtype=np.int32)
    out = np.zeros(ix.size, dtype=np.float64)
    lib().orc_fp_f64_rays(vol, W, H, float(sad), _angles(angles), geom, ix.size, ix, iy, view, out)
    return out


def fp_rays(vol, sad, angles, geom, ix, iy, view):
    """FP32 ray sums of the pixels (ix[r], iy[r]) of views view[r]."""
    vol = _f32(vol)
    W, H = vol.shape[0], vol.shape[2]
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    iy = np.ascontiguousarray(iy, dtype=np.int32)
    view = np.ascontiguousarray(view, dtype=np.int32)
    out = np.zeros(ix.size, dtype=np.float32)
    lib().orc_fp_rays(vol, W, H, float(sad), _angles(angles), geom, ix.size, ix, iy, view, out)
    return out


def bp_cols(sino, sad, angles, geom, ix, iy):
    """FP32 BP of voxel columns (ix[c], iy[c]) -> [ncol, H]."""
    sino = _f32(sino)
    W, H, n = sino.shape
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    iy = np.ascontiguousarray(iy, dtype=np.int32)
    out = np.zeros((H, ix.size), dtype=np.float32, order="F")
    lib().orc_bp_cols(sino, W, H, n, float(sad), _angles(angles), geom, ix.size, ix, iy, out)
    return out.T.copy()


def bp(sino, sad, angles, geom, f64=False):
    """CUDAmex_BP(P, SAD, angles, geomFlag) -> [W, W, H]."""
    sino = _f32(sino)
    W, H, n = sino.shape
    ang = _angles(angles)
    assert ang.size >= n
    if f64:
        out = np.zeros((W, W, H), dtype=np.float64, order="F")
        lib().orc_bp_f64(sino, W, H, n, float(sad), ang, geom, out)
    else:
        out = np.zeros((W, W, H), dtype=np.float32, order="F")
        lib().orc_bp(sino, W, H, n, float(sad), ang, geom, out)
    return out


def bp_f64_cols(sino, sad, angles, geom, ix, iy):
    sino = _f32(sino)
    W, H, n = sino.shape
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    iy = np.ascontiguousarray(iy, dtype=np.int32)
    out = np.zeros((H, ix.size), dtype=np.float64, order="F")
    lib().orc_bp_f64_cols(sino, W, H, n, float(sad), _angles(angles), geom, ix.size, ix, iy, out)
    return out.T.copy()


def g3d_fp(vol, angles, pts, nproj, f64=False):
    """CUDAmex_general3D_FP(vol, angles, pts, nProj); pts [W,H,6] or [W,H,6,nProj]."""
    vol = _f32(vol)
    pts = _f32(pts)
    W, H = pts.shape[0], pts.shape[1]
    ppv = 1 if pts.ndim == 4 else 0
    ang = _angles(angles)
    if f64:
        out = np.zeros((W, H, nproj), dtype=np.float64, order="F")
        lib().orc_g3d_fp_f64(vol, vol.shape[0], vol.shape[2], ang, pts, W, H, nproj, ppv, out)
    else:
        out = np.zeros((W, H, nproj), dtype=np.float32, order="F")
        lib().orc_g3d_fp(vol, vol.shape[0], vol.shape[2], ang, pts, W, H, nproj, ppv, out)
    return out


def g3d_bp(sino, angles, pts, M, N, f64=False):
    """CUDAmex_general3D_BP(P, angles, pts, M, N) -> [M, M, N]."""
    sino = _f32(sino)
    pts = _f32(pts)
    W, H, n = sino.shape
    ppv = 1 if pts.ndim == 4 else 0
    ang = _angles(angles)
    if f64:
        out = np.zeros((M, M, N), dtype=np.float64, order="F")
        lib().orc_g3d_bp_f64(sino, W, H, n, ang, pts, ppv, M, N, out)
    else:
        out = np.zeros((M, M, N), dtype=np.float32, order="F")
        lib().orc_g3d_bp(sino, W, H, n, ang, pts, ppv, M, N, out)
    return out


def osc_iter(mu, pre, post, num_subsets, picks, angles, sad, geom, f64=False):
    """CUDAmex_oscIter on copies; returns (mu', pre', post')."""
    dt = np.float64 if f64 else np.float32
    mu = np.array(mu, dtype=dt, order="F", copy=True)
    pre = np.array(pre, dtype=dt, order="F", copy=True)
    post = np.array(post, dtype=dt, order="F", copy=True)
    W, H = mu.shape[0], mu.shape[2]
    picks = np.ascontiguousarray(picks, dtype=np.int32)
    fn = lib().orc_osc_iter_f64 if f64 else lib().orc_osc_iter
    rc = fn(mu, W, H, pre, post, pre.shape[2], int(num_subsets), picks, _angles(angles),
            float(sad), geom)
    if rc:
        raise ValueError(f"orc_osc_iter failed ({rc})")
    return mu, pre, post


def tvmin(vol, c, steps=20, f64=False):
    dt = np.float64 if f64 else np.float32
    x = np.array(vol, dtype=dt, order="F", copy=True)
    fn = lib().orc_tvmin_f64 if f64 else lib().orc_tvmin
    fn(x, x.shape[0], x.shape[1], x.shape[2], float(c), int(steps))
    return x


def tv_gradient(vol):
    x = _f32(vol)
    g = np.zeros_like(x, order="F")
    lib().orc_tv_gradient(x, x.shape[0], x.shape[1], x.shape[2], g)
    return g


def design_filter(window, length, d=1.0):
    w = WINDOWS[window] if isinstance(window, str) else int(window)
    order = lib().orc_design_filter(w, int(length), float(d), None)
    out = np.zeros(order, dtype=np.float64)
    lib().orc_design_filter(w, int(length), float(d), out.ctypes.data)
    return out


GEOM_TYPES = {"par3d": 0, "fan3d": 1, "cone3d": 2, "cone": 2}


def _geom(geom):
    return GEOM_TYPES[geom] if isinstance(geom, str) else int(geom)


def prefilter(sino, geom, sad, angles, window="hamming", d=1.0, f64=False, parker_q=None):
    """preFilterSinogram(sino, geom, angles, filter, cutoff[, parker_q]);
    parker_q None is the 5-argument call (no Parker weighting)."""
    sino = _f32(sino)
    W, H, n = sino.shape
    g = _geom(geom)
    w = WINDOWS[window] if isinstance(window, str) else int(window)
    q = -1.0 if parker_q is None else float(parker_q)
    if f64:
        out = np.zeros((W, H, n), dtype=np.float64, order="F")
        rc = lib().orc_prefilter_f64(sino, W, H, n, g, float(sad), _angles(angles), w, float(d), q,
                                     out)
    else:
        out = np.zeros((W, H, n), dtype=np.float32, order="F")
        rc = lib().orc_prefilter(sino, W, H, n, g, float(sad), _angles(angles), w, float(d), q, out)
    if rc:
        raise ValueError(f"orc_prefilter failed ({rc})")
    return out


def parker_window(W, angles, geom, sad, q):
    """parker_weight.m:29-76 window, shape (W, nProj): column u of view i."""
    ang = _angles(angles)
    out = np.zeros((W, ang.size), dtype=np.float32, order="F")
    if lib().orc_parker_window(int(W), ang.size, _geom(geom), float(sad), ang, float(q), out):
        raise ValueError("orc_parker_window failed")
    return out


def parker_weight(sino, angles, geom, sad, q):
    """weightedSino = parker_weight(sino, angles, geom, q) (parker_weight.m:1-79)."""
    sino = _f32(sino)
    win = parker_window(sino.shape[0], angles, geom, sad, q)
    return np.asfortranarray(sino * win[:, None, :])


def log_transform(pre, post):
    """-log(post./pre) with NaN/Inf -> 0 (FBP_recon.m:26-28)."""
    pre = _f32(pre)
    post = _f32(post)
    out = np.zeros(pre.shape, dtype=np.float32, order="F")
    lib().orc_log_transform(pre, post, pre.size, out)
    return out


def fbp_recon(pre, post, angles, geom, sad, window="hamming", d=1.0, parker_q=None):
    """FBP_recon.m:26-38: log transform, pre-filter, then the BP of the
    geometry (cone for 'cone'/'cone3d')."""
    sino = log_transform(pre, post)
    filt = prefilter(sino, geom, sad, angles, window, d, parker_q=parker_q)
    return bp(filt, np.float32(sad), angles, _geom(geom))


def calc_num_subsets(s_init, s_final, p, K, k):
    return lib().orc_calc_num_subsets(int(s_init), int(s_final), float(p), int(K), int(k))


def calc_subset_order(n, first_pick):
    out = np.zeros(n, dtype=np.int32)
    lib().orc_calc_subset_order(int(n), int(first_pick), out)
    return out


def osc_tv_recon(pre, post, angles, sad, geom, n_iter, tv, s_init, s_final, p, tv_steps,
                 first_picks, f64=False):
    """OSC_TV_recon in float (the reference arithmetic) or, with f64=True, in
    double throughout."""
    pre = _f32(pre)
    post = _f32(post)
    W, H, n = pre.shape
    mu = np.zeros((W, W, H), dtype=np.float64 if f64 else np.float32, order="F")
    fp_ = np.ascontiguousarray(first_picks, dtype=np.int32)
    fn = lib().orc_osc_tv_recon_f64 if f64 else lib().orc_osc_tv_recon
    rc = fn(pre, post, W, H, n, _angles(angles), float(sad), int(geom), int(n_iter), float(tv),
            int(s_init), int(s_final), float(p), int(tv_steps), fp_, mu)
    if rc:
        raise ValueError(f"orc_osc_tv_recon failed ({rc})")
    return mu
