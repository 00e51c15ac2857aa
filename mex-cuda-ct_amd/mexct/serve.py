"""Continuous-batching request server (libmexct.so's mexct_server_*).

Many callers share one GPU: the server's workers (one HIP stream and
workspace each) take queued requests in arrival order while their device
footprints fit the HBM budget, and run same-shaped requests back to back.
Arrays are numpy, MATLAB layout (Fortran order), as in mexct.ops, or torch
tensors on the server's GPU in the layout of mexct.device (a volume
(H, W, W), a sinogram (nProj, H, W)): such requests are resident, their
results are tensors and no host copies are made.

    with Server(hbm_budget=64 << 30, slots=4) as s:
        t = s.submit_fdk(sino, {"type": "cone3d", "SAD": 1000.0}, angles)
        vol = s.result(t)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (EINVAL, F32, OP_BP, OP_FDK, OP_FP, OP_OSCTV, MexctError, Request, check,
                   lib)
from .ops import GEOM, WINDOWS, _dims3, _ptr, _single, _vec, first_picks


def _is_dev(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _dev3(t, name):
    """A resident request's float32 contiguous 3-D device tensor; the caller's
    stream is drained so the server's workers see its contents."""
    import torch
    if t.dtype != torch.float32 or t.dim() != 3 or not t.is_contiguous():
        raise MexctError(EINVAL, f"{name} must be a contiguous 3-D float32 device tensor")
    torch.cuda.current_stream(t.device).synchronize()
    return t

STAT_NAMES = ("done", "groups", "grouped", "queued", "reserved_bytes", "peak_reserved_bytes",
              "releases", "batched", "batches")


def request_footprint(op, W, H, nProj, geom=2, sad=1000.0, storage=F32, osc_iters=0,
                      s_init=1, s_final=1, p=1.0) -> int:
    """Bytes of device memory the server reserves for one request of this
    shape (mexct_server_footprint): its staging arrays plus the workspace
    of the entry point it runs.  op: OP_FP, OP_BP, OP_FDK or OP_OSCTV."""
    r = Request()
    r.op, r.geom, r.storage = int(op), int(geom), int(storage)
    r.W, r.H, r.nProj = int(W), int(H), int(nProj)
    r.sad = float(np.float32(sad))
    r.osc_iters, r.osc_s_init, r.osc_s_final, r.osc_p = int(osc_iters), int(s_init), int(s_final), float(p)
    b = int(lib().mexct_server_footprint(C.byref(r)))
    if b < 0:
        check(b)
    return b


class Server:
    def __init__(self, device: int = 0, hbm_budget: int = 32 << 30, slots: int = 4,
                 max_group: int = 4):
        self._h = C.c_void_p()
        check(lib().mexct_server_create(C.byref(self._h), int(device), int(hbm_budget),
                                        int(slots), int(max_group)))
        self._live: dict[int, tuple] = {}

    # ---------------------------------------------------------------- submit
    def _submit(self, req: Request, keep: tuple, out: np.ndarray) -> int:
        if not self._h:
            raise MexctError(EINVAL, "server is closed")
        t = C.c_uint64()
        check(lib().mexct_server_submit(self._h, C.byref(req), C.byref(t)))
        self._live[t.value] = (out, keep)
        return t.value

    @staticmethod
    def _base(op, geom, sad, angles, W, H, n, storage):
        ang = _vec(angles, np.float32, "angles")
        if ang.size < n:
            raise MexctError(EINVAL, "fewer angles than projections")
        r = Request()
        r.op, r.geom, r.storage = op, int(geom), int(storage)
        r.W, r.H, r.nProj = int(W), int(H), int(n)
        r.sad = float(np.float32(sad))
        r.angles = ang.ctypes.data
        return r, ang

    def submit_fp(self, vol, SAD, angles, nProjections, geomFlag, storage=F32) -> int:
        """CUDAmex_FP(vol, SAD, angles, nProjections, geomFlag) as a request."""
        if _is_dev(vol):
            import torch
            vol = _dev3(vol, "vol")
            H, W, W2 = vol.shape
            if W != W2:
                raise MexctError(EINVAL, "vol must be (H, W, W)")
            r, ang = self._base(OP_FP, geomFlag, SAD, angles, W, H, int(nProjections), storage)
            out = torch.empty((int(nProjections), H, W), dtype=torch.float32, device=vol.device)
            r.input, r.output, r.resident = vol.data_ptr(), out.data_ptr(), 1
            return self._submit(r, (vol, ang), out)
        vol = _single(vol, "vol")
        W, W2, H = _dims3(vol)
        if W != W2:
            raise MexctError(EINVAL, "vol must be [W,W,H]")
        r, ang = self._base(OP_FP, geomFlag, SAD, angles, W, H, int(nProjections), storage)
        out = np.zeros((W, H, int(nProjections)), np.float32, order="F")
        r.input, r.output = vol.ctypes.data, out.ctypes.data
        return self._submit(r, (vol, ang), out)

    def _submit_dev_sino(self, op, P, geom, SAD, angles, storage, window=0, cutoff=0.0):
        import torch
        P = _dev3(P, "sinogram")
        n, H, W = P.shape
        r, ang = self._base(op, geom, SAD, angles, W, H, n, storage)
        r.window, r.cutoff = int(window), float(cutoff)
        out = torch.empty((H, W, W), dtype=torch.float32, device=P.device)
        r.input, r.output, r.resident = P.data_ptr(), out.data_ptr(), 1
        return self._submit(r, (P, ang), out)

    def submit_bp(self, P, SAD, angles, geomFlag, storage=F32) -> int:
        """CUDAmex_BP(P, SAD, angles, geomFlag) as a request."""
        if _is_dev(P):
            return self._submit_dev_sino(OP_BP, P, geomFlag, SAD, angles, storage)
        P = _single(P, "P")
        W, H, n = _dims3(P)
        r, ang = self._base(OP_BP, geomFlag, SAD, angles, W, H, n, storage)
        out = np.zeros((W, W, H), np.float32, order="F")
        r.input, r.output = P.ctypes.data, out.ctypes.data
        return self._submit(r, (P, ang), out)

    def submit_fdk(self, sino, geom, angles, filter="hamming", frequencyCutoff=1.0,
                   storage=F32) -> int:
        """CUDAmex_BP(preFilterSinogram(sino, geom, angles, filter, cutoff))
        as one request (SAD is taken as single precision for both steps)."""
        if _is_dev(sino):
            return self._submit_dev_sino(OP_FDK, sino, GEOM[str(geom["type"]).lower()],
                                         geom.get("SAD", 0.0), angles, storage,
                                         WINDOWS[str(filter).lower()], float(frequencyCutoff))
        sino = _single(sino, "sino")
        W, H, n = _dims3(sino)
        g = GEOM[str(geom["type"]).lower()]
        r, ang = self._base(OP_FDK, g, geom.get("SAD", 0.0), angles, W, H, n, storage)
        r.window, r.cutoff = WINDOWS[str(filter).lower()], float(frequencyCutoff)
        out = np.zeros((W, W, H), np.float32, order="F")
        r.input, r.output = sino.ctypes.data, out.ctypes.data
        return self._submit(r, (sino, ang), out)

    def submit_osc_tv(self, pre, post, scanAngles, geom, numIter, TV_constant, numSubsInit,
                      numSubsFinal, pVal, tv_steps=20, seed=0, picks=None, storage=F32) -> int:
        """OSC_TV_recon(pre, post, angles, geom, nIter, TV, sInit, sFinal, p)
        as a request."""
        dev = _is_dev(pre)
        if dev:
            import torch
            pre, post = _dev3(pre, "pre"), _dev3(post, "post")
            n, H, W = pre.shape
            if tuple(post.shape) != (n, H, W) or post.device != pre.device:
                raise MexctError(EINVAL, "pre and post differ in shape or device")
        else:
            pre = _single(pre, "pre")
            post = _single(post, "post")
            W, H, n = _dims3(pre)
            if _dims3(post) != (W, H, n):
                raise MexctError(EINVAL, "pre and post differ in shape")
        g = GEOM[str(geom["type"]).lower()]
        r, ang = self._base(OP_OSCTV, g, geom["SAD"], scanAngles, W, H, n, storage)
        if picks is None:
            picks = first_picks(seed, numSubsInit, numSubsFinal, pVal, int(numIter))
        picks = _vec(picks, np.int32, "picks")
        if picks.size < int(numIter):
            raise MexctError(EINVAL, "picks needs numIter entries")
        r.osc_iters, r.osc_tv = int(numIter), float(np.float32(TV_constant))
        r.osc_s_init, r.osc_s_final, r.osc_p = int(numSubsInit), int(numSubsFinal), float(pVal)
        r.tv_steps, r.first_picks = int(tv_steps), picks.ctypes.data
        if dev:
            out = torch.empty((H, W, W), dtype=torch.float32, device=pre.device)
            r.input, r.input2, r.output = pre.data_ptr(), post.data_ptr(), out.data_ptr()
            r.resident = 1
        else:
            out = np.zeros((W, W, H), np.float32, order="F")
            r.input, r.input2, r.output = pre.ctypes.data, post.ctypes.data, out.ctypes.data
        return self._submit(r, (pre, post, ang, picks), out)

    # ---------------------------------------------------------------- results
    def wait(self, ticket: int):
        """(output, latency_s, queue_s); raises MexctError if the request failed."""
        st, lat, q = C.c_int32(), C.c_double(), C.c_double()
        check(lib().mexct_server_wait(self._h, C.c_uint64(ticket), C.byref(st), C.byref(lat),
                                      C.byref(q)))
        out, _ = self._live.pop(ticket)
        check(st.value)
        return out, lat.value, q.value

    def result(self, ticket: int) -> np.ndarray:
        return self.wait(ticket)[0]

    def stats(self) -> dict:
        v = (C.c_int64 * len(STAT_NAMES))()
        check(lib().mexct_server_stats(self._h, v, len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, list(v)))

    def close(self) -> None:
        if self._h:
            # waits for running requests; queued ones are dropped
            check(lib().mexct_server_destroy(self._h))
            self._h = C.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
