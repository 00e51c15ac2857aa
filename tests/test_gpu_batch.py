"""Batched launches and the block-order autotuner.

A batched FP or BP computes every item exactly as the single call does, so
the results are compared bit for bit; so are tuned and untuned runs, whose
candidates differ only in which workgroup computes which rays or voxels.
"""
import numpy as np
import pytest

import mexct
from mexct import device

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _angles(n, off=0.05):
    return (np.arange(n) * (2 * np.pi / n) + off).astype(np.float32)


@pytest.fixture(scope="module")
def torch_cuda(gpu):
    import torch
    return torch


CASES = [(40, 24, 18, 0, 500.0), (56, 33, 20, 1, 150.0), (64, 70, 24, 2, 300.0),
         (72, 40, 16, 2, 60.0)]


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
@pytest.mark.parametrize("case", CASES)
def test_fp_batch_equals_single(torch_cuda, case, storage):
    torch = torch_cuda
    W, H, n, geom, sad = case
    G = 3
    ang = _angles(n)
    vols = torch.from_numpy(_rng(W + geom).uniform(0, 1, (G, H, W, W)).astype(np.float32)).cuda()
    ctx = device.Context(0)
    outs = torch.zeros((G, n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp_batch(vols, sad, ang, n, geom, outs, storage=storage)
    one = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    for g in range(G):
        ctx.fp(vols[g].contiguous(), sad, ang, n, geom, one, storage=storage)
        torch.cuda.synchronize()
        assert torch.equal(outs[g], one), f"item {g} differs"
    ctx.close()


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
@pytest.mark.parametrize("case", CASES)
def test_bp_batch_equals_single(torch_cuda, case, storage):
    torch = torch_cuda
    W, H, n, geom, sad = case
    G = 3
    ang = _angles(n)
    sinos = torch.from_numpy(_rng(W * 3 + geom).uniform(-1, 1, (G, n, H, W)).astype(np.float32)).cuda()
    ctx = device.Context(0)
    outs = torch.zeros((G, H, W, W), dtype=torch.float32, device="cuda")
    ctx.bp_batch(sinos, sad, ang, geom, outs, storage=storage)
    one = torch.zeros((H, W, W), dtype=torch.float32, device="cuda")
    for g in range(G):
        ctx.bp(sinos[g].contiguous(), sad, ang, geom, one, storage=storage)
        torch.cuda.synchronize()
        assert torch.equal(outs[g], one), f"item {g} differs"
    ctx.close()


def test_batch_rejects_bad_counts(torch_cuda):
    torch = torch_cuda
    ctx = device.Context(0)
    v = torch.zeros((1, 8, 8, 8), dtype=torch.float32, device="cuda")
    o = torch.zeros((1, 4, 8, 8), dtype=torch.float32, device="cuda")
    from mexct import _lib as L
    import ctypes as C
    ang = np.zeros(4, np.float32)
    rc = L.lib().mexct_fp_batch_dev(ctx._h, C.c_void_p(v.data_ptr()), 0, 0, 8, 8, C.c_float(100.0),
                                    ang.ctypes.data, 4, 2, C.c_void_p(o.data_ptr()), None)
    assert rc == L.EINVAL
    rc = L.lib().mexct_bp_batch_dev(ctx._h, C.c_void_p(o.data_ptr()), 0, 257, 8, 8, 4,
                                    C.c_float(100.0), ang.ctypes.data, 2, C.c_void_p(v.data_ptr()),
                                    None)
    assert rc == L.EINVAL
    ctx.close()


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
def test_autotune_same_bits_and_remembered(torch_cuda, storage):
    torch = torch_cuda
    W, H, n, sad = 88, 52, 36, 400.0  # a shape no other test uses
    ang = _angles(n)
    r = _rng(90 + storage)
    vol = torch.from_numpy(r.uniform(0, 1, (H, W, W)).astype(np.float32)).cuda()
    sino = torch.from_numpy(r.uniform(-1, 1, (n, H, W)).astype(np.float32)).cuda()
    plain = device.Context(0)
    plain.set_autotune(0)
    fp0 = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    bp0 = torch.zeros((H, W, W), dtype=torch.float32, device="cuda")
    plain.fp(vol, sad, ang, n, 2, fp0, storage=storage)
    plain.bp(sino, sad, ang, 2, bp0, storage=storage)
    assert mexct.tuned_config(0, 2, storage, W, H, n)[0] == 0
    tuned = device.Context(0)
    tuned.set_autotune(1)
    fp1 = torch.zeros_like(fp0)
    bp1 = torch.zeros_like(bp0)
    tuned.fp(vol, sad, ang, n, 2, fp1, storage=storage)   # tunes
    tuned.bp(sino, sad, ang, 2, bp1, storage=storage)
    torch.cuda.synchronize()
    assert torch.equal(fp0, fp1) and torch.equal(bp0, bp1)
    cf = mexct.tuned_config(0, 2, storage, W, H, n)
    cb = mexct.tuned_config(1, 2, storage, W, H, n)
    print(f"tuned FP {cf}, BP {cb}")
    assert cf[0] == 1 and 2 <= cf[1] <= 5
    assert cb[0] == 1 and cb[1] in (2, 4, 8) and cb[2] in (32, 64)
    # later calls of the shape, from any context, use the stored choice
    fp2 = torch.zeros_like(fp0)
    bp2 = torch.zeros_like(bp0)
    plain.fp(vol, sad, ang, n, 2, fp2, storage=storage)
    plain.bp(sino, sad, ang, 2, bp2, storage=storage)
    torch.cuda.synchronize()
    assert torch.equal(fp0, fp2) and torch.equal(bp0, bp2)
    plain.close()
    tuned.close()


def test_autotune_mode2_tunes_recurring_shapes(torch_cuda):
    torch = torch_cuda
    W, H, n, sad = 76, 30, 22, 250.0  # a shape no other test uses
    ang = _angles(n)
    vol = torch.from_numpy(_rng(95).uniform(0, 1, (H, W, W)).astype(np.float32)).cuda()
    ctx = device.Context(0)
    ctx.set_autotune(2)
    outs = [torch.zeros((n, H, W), dtype=torch.float32, device="cuda") for _ in range(3)]
    ctx.fp(vol, sad, ang, n, 1, outs[0])
    assert mexct.tuned_config(0, 1, mexct.F32, W, H, n)[0] == 0  # one-off so far
    ctx.fp(vol, sad, ang, n, 1, outs[1])
    assert mexct.tuned_config(0, 1, mexct.F32, W, H, n)[0] == 1  # tuned at the 2nd call
    ctx.fp(vol, sad, ang, n, 1, outs[2])
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
    ctx.close()


def test_tuned_shape_keeps_a_fixed_zrange(torch_cuda):
    # 2-byte storage steps the rows in texel space from the z-range start,
    # so its bits depend on the z-range: a shape tuned by a caller that
    # fixed zrange = 64 must not hand 64 to a later caller that asked for
    # the default, and a later zrange = 64 call keeps 64
    torch = torch_cuda
    W, H, n, sad = 84, 100, 20, 350.0  # a shape no other test uses
    ang = _angles(n)
    sino = torch.from_numpy(_rng(97).uniform(-1, 1, (n, H, W)).astype(np.float32)).cuda()
    plain = device.Context(0)
    plain.set_autotune(0)
    want = {}
    for zr in (0, 64):
        want[zr] = torch.zeros((H, W, W), dtype=torch.float32, device="cuda")
        plain.bp(sino, sad, ang, 2, want[zr], storage=mexct.F16, zrange=zr)
    torch.cuda.synchronize()
    assert not torch.equal(want[0], want[64])  # the range moves the bits
    tuned = device.Context(0)
    tuned.set_autotune(1)
    got = torch.zeros_like(want[0])
    tuned.bp(sino, sad, ang, 2, got, storage=mexct.F16, zrange=64)  # tunes the shape
    torch.cuda.synchronize()
    assert mexct.tuned_config(1, 2, mexct.F16, W, H, n)[0] == 1
    assert torch.equal(got, want[64])
    for zr in (0, 64):
        plain.bp(sino, sad, ang, 2, got, storage=mexct.F16, zrange=zr)
        torch.cuda.synchronize()
        assert torch.equal(got, want[zr]), f"zrange {zr} after tuning"
    plain.close()
    tuned.close()
