"""Device-resident use of the operators: a Context owns a HIP stream and a
workspace on one GPU, and its methods take device tensors (torch is used
only to hold device memory and to name the stream).

Shapes are MATLAB's dims in Fortran order.  A torch tensor is row-major,
so a [W, W, H] volume is held as a contiguous tensor of shape (H, W, W)
(index [z, y, x]) and a [W, H, nProj] sinogram as (nProj, H, W).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BF16, F16, F32, EINVAL, MexctError, check, lib, opts

_DT = {"torch.float32": F32, "torch.float16": F16, "torch.bfloat16": BF16}


def _dtype_code(t):
    code = _DT.get(str(t.dtype))
    if code is None:
        raise MexctError(EINVAL, f"unsupported tensor dtype {t.dtype}")
    return code


def _dptr(t, name, want_f32=False):
    if not t.is_cuda or not t.is_contiguous():
        raise MexctError(EINVAL, f"{name} must be a contiguous device tensor")
    if want_f32 and str(t.dtype) != "torch.float32":
        raise MexctError(EINVAL, f"{name} must be float32")
    return C.c_void_p(t.data_ptr())


def _host_f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


class Context:
    """mexct_ctx on `device`; stream=None uses torch's current stream there."""

    def __init__(self, device: int = 0, stream=None):
        import torch

        self.device = int(device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        h = C.c_void_p()
        check(lib().mexct_ctx_create(C.byref(h), self.device, C.c_void_p(int(stream))))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().mexct_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().mexct_ctx_sync(self._h))

    def workspace_bytes(self) -> int:
        return int(lib().mexct_ctx_workspace_bytes(self._h))

    # ------------------------------------------------------------ operators
    def fp(self, vol, sad, angles, nproj, geom, out, storage=F32):
        """vol (H, W, W) any of f32/f16/bf16 -> out (nProj, H, W) f32."""
        H, W, W2 = vol.shape
        ang = _host_f32(angles)
        if W2 != W or ang.size < nproj or tuple(out.shape) != (nproj, H, W):
            raise MexctError(EINVAL, "fp: shape mismatch")
        check(lib().mexct_fp_dev(self._h, _dptr(vol, "vol"), _dtype_code(vol), W, H, float(sad),
                                 ang.ctypes.data_as(C.c_void_p), int(nproj), int(geom),
                                 _dptr(out, "out", True), opts(storage)))
        return out

    def bp(self, sino, sad, angles, geom, out, storage=F32, zrange=0):
        """sino (nProj, H, W) -> out (H, W, W) f32."""
        n, H, W = sino.shape
        ang = _host_f32(angles)
        if ang.size < n or tuple(out.shape) != (H, W, W):
            raise MexctError(EINVAL, "bp: shape mismatch")
        check(lib().mexct_bp_dev(self._h, _dptr(sino, "sino"), _dtype_code(sino), W, H, n,
                                 float(sad), ang.ctypes.data_as(C.c_void_p), int(geom),
                                 _dptr(out, "out", True), opts(storage, 0, zrange)))
        return out

    def g3d_fp(self, vol, angles, pts, nproj, out, storage=F32):
        """pts (6, H, W) shared or (nProj, 6, H, W) per view."""
        nsl, side, _ = vol.shape
        ppv = 1 if pts.dim() == 4 else 0
        H, W = pts.shape[-2], pts.shape[-1]
        ang = _host_f32(angles)
        check(lib().mexct_g3d_fp_dev(self._h, _dptr(vol, "vol"), _dtype_code(vol), side, nsl,
                                     ang.ctypes.data_as(C.c_void_p), _dptr(pts, "pts", True), W,
                                     H, int(nproj), ppv, _dptr(out, "out", True), opts(storage)))
        return out

    def g3d_bp(self, sino, angles, pts, M, N, out):
        n, H, W = sino.shape
        ppv = 1 if pts.dim() == 4 else 0
        ang = _host_f32(angles)
        check(lib().mexct_g3d_bp_dev(self._h, _dptr(sino, "sino", True), W, H, n,
                                     ang.ctypes.data_as(C.c_void_p), _dptr(pts, "pts", True), ppv,
                                     int(M), int(N), _dptr(out, "out", True), None))
        return out

    def osc_iter(self, mu, pre, post, num_subsets, picks, angles, sad, geom):
        H, W, _ = mu.shape
        n = pre.shape[0]
        pk = np.ascontiguousarray(np.asarray(picks, dtype=np.int32).ravel())
        ang = _host_f32(angles)
        check(lib().mexct_osc_iter_dev(self._h, _dptr(mu, "mu", True), W, H,
                                       _dptr(pre, "pre", True), _dptr(post, "post", True), n,
                                       int(num_subsets), pk.ctypes.data_as(C.c_void_p),
                                       ang.ctypes.data_as(C.c_void_p), float(sad), int(geom), None))

    def tvmin(self, vol, c, steps=20):
        Nz, Ny, Nx = vol.shape
        check(lib().mexct_tvmin_dev(self._h, _dptr(vol, "vol", True), Nx, Ny, Nz, float(c),
                                    opts(F32, int(steps))))
        return vol

    def prefilter(self, sino, geom, sad, angles, window, cutoff, out):
        n, H, W = sino.shape
        ang = _host_f32(angles)
        check(lib().mexct_prefilter_dev(self._h, _dptr(sino, "sino", True), W, H, n, int(geom),
                                        float(sad), ang.ctypes.data_as(C.c_void_p), int(window),
                                        float(cutoff), _dptr(out, "out", True)))
        return out

    def osc_tv_recon(self, pre, post, angles, sad, geom, n_iter, tv, s_init, s_final, p,
                     first_picks, out, tv_steps=20):
        n, H, W = pre.shape
        ang = _host_f32(angles)
        fp_ = np.ascontiguousarray(np.asarray(first_picks, dtype=np.int32).ravel())
        check(lib().mexct_osc_tv_recon_dev(self._h, _dptr(pre, "pre", True),
                                           _dptr(post, "post", True), W, H, n,
                                           ang.ctypes.data_as(C.c_void_p), float(sad), int(geom),
                                           int(n_iter), float(tv), int(s_init), int(s_final),
                                           float(p), fp_.ctypes.data_as(C.c_void_p),
                                           _dptr(out, "out", True), opts(F32, int(tv_steps))))
        return out
