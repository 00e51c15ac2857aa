"""The allocation-free pre-filter (mexct_prefilter_noalloc_dev) against the
pre-filter it mirrors (mexct_prefilter_dev), through the C ABI.

Both run the same kernel on the same tables, so the outputs must be equal
bit for bit.  The calls go in a row, the second smaller than the first, so
that the second reuses the fixed scratch the first filled; then a repeat
of a call (its tables are kept) and a parallel-beam call with a Parker
window."""
import ctypes as C

import numpy as np
import pytest

from mexct import _lib as L
from mexct import device

pytestmark = pytest.mark.gpu

# (W, H, nProj, geom, sad, window, cutoff, parker_q, angular span)
CALLS = [
    (256, 200, 180, 2, 700.0, 3, 1.0, 0.5, np.pi * 1.25),
    (96, 64, 90, 1, 300.0, 1, 0.8, None, 2 * np.pi),
    (96, 64, 90, 1, 300.0, 1, 0.8, None, 2 * np.pi),
    (64, 40, 120, 0, 1000.0, 4, 1.0, 1.0, np.pi),
]


def _noalloc(ctx, sino, W, H, n, geom, sad, ang, window, cutoff, q, out):
    import torch

    torch.cuda.synchronize()
    rc = L.lib().mexct_prefilter_noalloc_dev(
        ctx._h, C.c_void_p(sino.data_ptr()), W, H, n, geom, float(sad),
        ang.ctypes.data_as(C.c_void_p), window, float(cutoff), float(q),
        C.c_void_p(out.data_ptr()))
    L.check(rc)
    ctx.sync()


def test_noalloc_prefilter_equals_prefilter(gpu):
    import torch

    ctx = device.Context(0)
    try:
        for i, (W, H, n, geom, sad, window, cutoff, q, span) in enumerate(CALLS):
            print(f"call {i}: W={W} H={H} nProj={n} geom={geom} parker={q}", flush=True)
            r = np.random.Generator(np.random.PCG64(100 + i))
            sino = torch.from_numpy(r.uniform(0.0, 2.0, (n, H, W)).astype(np.float32)).cuda()
            ang = np.ascontiguousarray((np.arange(n) * span / n).astype(np.float32))
            want = torch.full((n, H, W), np.nan, dtype=torch.float32, device="cuda")
            ctx.prefilter(sino, geom, sad, ang, window, cutoff, want, parker_q=q)
            ctx.sync()
            got = torch.full((n, H, W), np.nan, dtype=torch.float32, device="cuda")
            _noalloc(ctx, sino, W, H, n, geom, sad, ang, window, cutoff,
                     -1.0 if q is None else q, got)
            assert torch.isfinite(want).all()
            assert torch.equal(got, want), f"call {i} differs"
    finally:
        ctx.close()


def test_noalloc_prefilter_rejects_tables_beyond_scratch(gpu):
    import torch

    ctx = device.Context(0)
    try:
        W = H = 2048  # the cone weights alone fill the 16 MiB
        x = torch.zeros(1, device="cuda")
        ang = np.zeros(4, np.float32)
        rc = L.lib().mexct_prefilter_noalloc_dev(ctx._h, C.c_void_p(x.data_ptr()), W, H, 4, 2,
                                                 1000.0, ang.ctypes.data_as(C.c_void_p), 3, 1.0,
                                                 -1.0, C.c_void_p(x.data_ptr()))
        assert rc == L.EINVAL
        assert b"allocation-free" in L.lib().mexct_last_error()
    finally:
        ctx.close()
