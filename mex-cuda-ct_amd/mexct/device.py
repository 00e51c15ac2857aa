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


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _host_f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


def _need(ok, what):
    """Raise EINVAL unless ok: sizes are checked here because the _dev entry
    points take device pointers without lengths."""
    if not ok:
        raise MexctError(EINVAL, what)


def _shape(t, want, name):
    _need(tuple(t.shape) == tuple(int(x) for x in want),
          f"{name} has shape {tuple(t.shape)}, expected {tuple(int(x) for x in want)}")


class Context:
    """mexct_ctx on `device` running on its own HIP stream.

    stream: None creates a torch.cuda.Stream (default priority); a
    torch.cuda.Stream or a raw hipStream_t handle (int) is used as given.
    order_with_torch: each call first waits for torch's current stream and
    makes torch's current stream wait for it afterwards, so tensors produced
    or consumed by torch work are ordered with the operators."""

    def __init__(self, device: int = 0, stream=None, order_with_torch: bool = True):
        import torch

        self.device = int(device)
        if stream is None:
            stream = torch.cuda.Stream(device=self.device)
        self.stream = stream if isinstance(stream, torch.cuda.Stream) else None
        handle = stream.cuda_stream if isinstance(stream, torch.cuda.Stream) else int(stream)
        if handle == 0:
            raise MexctError(EINVAL, "pass a non-null stream (the null stream cannot be shared)")
        self.order_with_torch = order_with_torch and self.stream is not None
        h = C.c_void_p()
        check(lib().mexct_ctx_create(C.byref(h), self.device, C.c_void_p(handle)))
        self._h = h

    def _before(self):
        if self.order_with_torch:
            import torch
            self.stream.wait_stream(torch.cuda.current_stream(self.device))

    def _after(self):
        if self.order_with_torch:
            import torch
            torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def _call(self, fn, *args):
        self._before()
        rc = fn(*args)
        self._after()
        check(rc)

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

    def probe_gather(self, storage, W, H, nviews, pattern=0):
        """Gather microbenchmark launch (see mexct_probe_gather_dev): pattern
        0 is the forward projector's loads, 1 the backprojector's, plus k << 4
        for the loop unrolled 2^k times (k = 0..3); returns the sample
        (update) count of the launch."""
        n = C.c_int64()
        self._call(lib().mexct_probe_gather_dev, self._h, int(storage), int(W), int(H),
                   int(nviews), int(pattern), C.byref(n))
        return n.value

    # ------------------------------------------------------------ operators
    def fp(self, vol, sad, angles, nproj, geom, out, storage=F32):
        """vol (H, W, W) any of f32/f16/bf16 -> out (nProj, H, W) f32."""
        H, W, W2 = vol.shape
        ang = _host_f32(angles)
        if W2 != W or ang.size < nproj or tuple(out.shape) != (nproj, H, W):
            raise MexctError(EINVAL, "fp: shape mismatch")
        self._call(lib().mexct_fp_dev, self._h, _dptr(vol, "vol"), _dtype_code(vol), W, H,
                   float(sad), _p(ang), int(nproj), int(geom), _dptr(out, "out", True),
                   opts(storage))
        return out

    def fp_loads(self, vol, sad, angles, nproj, geom, out, storage=F32):
        """The forward projector's loads-only form (mexct_fp_loads_dev): the
        same launch without the interpolation, its gather ceiling."""
        H, W, W2 = vol.shape
        ang = _host_f32(angles)
        if W2 != W or ang.size < nproj or tuple(out.shape) != (nproj, H, W):
            raise MexctError(EINVAL, "fp_loads: shape mismatch")
        self._call(lib().mexct_fp_loads_dev, self._h, _dptr(vol, "vol"), _dtype_code(vol), W, H,
                   float(sad), _p(ang), int(nproj), int(geom), _dptr(out, "out", True),
                   opts(storage))
        return out

    def bp(self, sino, sad, angles, geom, out, storage=F32, zrange=0):
        """sino (nProj, H, W) -> out (H, W, W) f32."""
        n, H, W = sino.shape
        ang = _host_f32(angles)
        if ang.size < n or tuple(out.shape) != (H, W, W):
            raise MexctError(EINVAL, "bp: shape mismatch")
        self._call(lib().mexct_bp_dev, self._h, _dptr(sino, "sino"), _dtype_code(sino), W, H, n,
                   float(sad), _p(ang), int(geom), _dptr(out, "out", True),
                   opts(storage, 0, zrange))
        return out

    def bp_loads(self, sino, sad, angles, geom, out, storage=F32, zrange=0):
        """The backprojector's loads-only form (mexct_bp_loads_dev, cone)."""
        n, H, W = sino.shape
        ang = _host_f32(angles)
        if ang.size < n or tuple(out.shape) != (H, W, W):
            raise MexctError(EINVAL, "bp_loads: shape mismatch")
        self._call(lib().mexct_bp_loads_dev, self._h, _dptr(sino, "sino"), _dtype_code(sino), W,
                   H, n, float(sad), _p(ang), int(geom), _dptr(out, "out", True),
                   opts(storage, 0, zrange))
        return out

    def fp_batch(self, vols, sad, angles, nproj, geom, outs, storage=F32):
        """vols (G, H, W, W) -> outs (G, nProj, H, W) f32, one launch."""
        G, H, W, W2 = vols.shape
        ang = _host_f32(angles)
        if W2 != W or ang.size < nproj or tuple(outs.shape) != (G, nproj, H, W):
            raise MexctError(EINVAL, "fp_batch: shape mismatch")
        self._call(lib().mexct_fp_batch_dev, self._h, _dptr(vols, "vols"), _dtype_code(vols), G,
                   W, H, float(sad), _p(ang), int(nproj), int(geom), _dptr(outs, "outs", True),
                   opts(storage))
        return outs

    def bp_batch(self, sinos, sad, angles, geom, outs, storage=F32, zrange=0):
        """sinos (G, nProj, H, W) -> outs (G, H, W, W) f32, one launch per view chunk."""
        G, n, H, W = sinos.shape
        ang = _host_f32(angles)
        if ang.size < n or tuple(outs.shape) != (G, H, W, W):
            raise MexctError(EINVAL, "bp_batch: shape mismatch")
        self._call(lib().mexct_bp_batch_dev, self._h, _dptr(sinos, "sinos"), _dtype_code(sinos), G,
                   W, H, n, float(sad), _p(ang), int(geom), _dptr(outs, "outs", True),
                   opts(storage, 0, zrange))
        return outs

    def set_autotune(self, mode):
        """-1: follow mexct.set_autotune; 0 off; 1 tune a problem shape at its
        first call; 2 at its second."""
        check(lib().mexct_ctx_set_autotune(self._h, int(mode)))

    def g3d_fp(self, vol, angles, pts, nproj, out, storage=F32):
        """pts (6, H, W) shared or (nProj, 6, H, W) per view."""
        _need(vol.dim() == 3 and vol.shape[1] == vol.shape[2], "g3d_fp: vol must be (nSlices, side, side)")
        nsl, side, _ = vol.shape
        _need(pts.dim() in (3, 4) and pts.shape[-3] == 6, "g3d_fp: pts must be (6,H,W) or (nProj,6,H,W)")
        ppv = 1 if pts.dim() == 4 else 0
        H, W = pts.shape[-2], pts.shape[-1]
        nproj = int(nproj)
        _need(nproj >= 0 and (not ppv or pts.shape[0] >= nproj), "g3d_fp: per-view pts has fewer views than nProj")
        ang = _host_f32(angles)
        _need(ang.size >= nproj, "g3d_fp: fewer angles than nProj")
        _shape(out, (nproj, H, W), "g3d_fp: out")
        self._call(lib().mexct_g3d_fp_dev, self._h, _dptr(vol, "vol"), _dtype_code(vol), side, nsl,
                   _p(ang), _dptr(pts, "pts", True), W, H, int(nproj), ppv,
                   _dptr(out, "out", True), opts(storage))
        return out

    def g3d_bp(self, sino, angles, pts, M, N, out, mode=0):
        """mode: 0 voxel bricks, 1 deterministic bricks, 2 ray-driven atomics."""
        n, H, W = sino.shape
        _need(pts.dim() in (3, 4) and tuple(pts.shape[-3:]) == (6, H, W),
              "g3d_bp: pts must be (6,H,W) or (nProj,6,H,W) matching sino")
        ppv = 1 if pts.dim() == 4 else 0
        _need(not ppv or pts.shape[0] >= n, "g3d_bp: per-view pts has fewer views than the sinogram")
        ang = _host_f32(angles)
        _need(ang.size >= n, "g3d_bp: fewer angles than views")
        _shape(out, (int(N), int(M), int(M)), "g3d_bp: out")
        self._call(lib().mexct_g3d_bp_dev, self._h, _dptr(sino, "sino", True), W, H, n, _p(ang),
                   _dptr(pts, "pts", True), ppv, int(M), int(N), _dptr(out, "out", True),
                   opts(g3d_bp=int(mode)))
        return out

    def osc_iter(self, mu, pre, post, num_subsets, picks, angles, sad, geom, storage=F32):
        _need(mu.dim() == 3 and mu.shape[1] == mu.shape[2], "osc_iter: mu must be (H, W, W)")
        H, W, _ = mu.shape
        _need(pre.dim() == 3 and tuple(pre.shape[1:]) == (H, W), "osc_iter: pre must be (nProj, H, W)")
        n = pre.shape[0]
        _shape(post, pre.shape, "osc_iter: post")
        pk = np.ascontiguousarray(np.asarray(picks, dtype=np.int32).ravel())
        _need(pk.size >= int(num_subsets), "osc_iter: fewer subset picks than subsets")
        ang = _host_f32(angles)
        _need(ang.size >= n, "osc_iter: fewer angles than views")
        self._call(lib().mexct_osc_iter_dev, self._h, _dptr(mu, "mu", True), W, H,
                   _dptr(pre, "pre", True), _dptr(post, "post", True), n, int(num_subsets),
                   _p(pk), _p(ang), float(sad), int(geom), opts(int(storage)))

    def tvmin(self, vol, c, steps=20):
        _need(vol.dim() == 3, "tvmin: vol must be (Nz, Ny, Nx)")
        Nz, Ny, Nx = vol.shape
        self._call(lib().mexct_tvmin_dev, self._h, _dptr(vol, "vol", True), Nx, Ny, Nz, float(c),
                   opts(F32, int(steps)))
        return vol

    def prefilter(self, sino, geom, sad, angles, window, cutoff, out, parker_q=None):
        n, H, W = sino.shape
        ang = _host_f32(angles)
        _need(ang.size >= n, "prefilter: fewer angles than views")
        _shape(out, sino.shape, "prefilter: out")
        q = -1.0 if parker_q is None else float(parker_q)
        self._call(lib().mexct_prefilter_dev, self._h, _dptr(sino, "sino", True), W, H, n,
                   int(geom), float(sad), _p(ang), int(window), float(cutoff), q,
                   _dptr(out, "out", True))
        return out

    def parker_weight(self, sino, angles, geom, sad, q, out):
        n, H, W = sino.shape
        ang = _host_f32(angles)
        _need(ang.size >= n, "parker_weight: fewer angles than views")
        _shape(out, sino.shape, "parker_weight: out")
        self._call(lib().mexct_parker_weight_dev, self._h, _dptr(sino, "sino", True), W, H, n,
                   _p(ang), int(geom), float(sad), float(q), _dptr(out, "out", True))
        return out

    def fbp_recon(self, pre, post, angles, geom, sad, window, cutoff, out, parker_q=None,
                  storage=F32):
        n, H, W = pre.shape
        ang = _host_f32(angles)
        _need(ang.size >= n, "fbp_recon: fewer angles than views")
        _shape(post, pre.shape, "fbp_recon: post")
        _shape(out, (H, W, W), "fbp_recon: out")
        q = -1.0 if parker_q is None else float(parker_q)
        self._call(lib().mexct_fbp_recon_dev, self._h, _dptr(pre, "pre", True),
                   _dptr(post, "post", True), W, H, n, _p(ang), int(geom), float(sad),
                   int(window), float(cutoff), q, _dptr(out, "out", True), opts(int(storage)))
        return out

    def osc_tv_recon(self, pre, post, angles, sad, geom, n_iter, tv, s_init, s_final, p,
                     first_picks, out, tv_steps=20, storage=F32):
        n, H, W = pre.shape
        ang = _host_f32(angles)
        fp_ = np.ascontiguousarray(np.asarray(first_picks, dtype=np.int32).ravel())
        _need(ang.size >= n, "osc_tv_recon: fewer angles than views")
        _need(fp_.size >= int(n_iter), "osc_tv_recon: fewer first picks than iterations")
        _shape(post, pre.shape, "osc_tv_recon: post")
        _shape(out, (H, W, W), "osc_tv_recon: out")
        self._call(lib().mexct_osc_tv_recon_dev, self._h, _dptr(pre, "pre", True),
                   _dptr(post, "post", True), W, H, n, _p(ang), float(sad), int(geom), int(n_iter),
                   float(tv), int(s_init), int(s_final), float(p), _p(fp_),
                   _dptr(out, "out", True), opts(int(storage), int(tv_steps)))
        return out
