"""Edge cases of the FP/BP entry points against the oracle: no views, the
smallest volumes, odd sizes, repeated and negative angles, a single slice,
rays that miss the volume (small SAD), and the TV step on a flat volume."""
import numpy as np
import pytest

import mexct
import oracle as O

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _vol(W, H, seed):
    return np.asfortranarray(_rng(seed).uniform(0, 1, (W, W, H)).astype(np.float32))


@pytest.mark.parametrize("geom", [0, 1, 2])
def test_no_views(gpu, geom):
    vol = _vol(8, 6, 1)
    fp = mexct.CUDAmex_FP(vol, 100.0, np.zeros(3, np.float32), 0, geom)
    assert fp.shape == (8, 6, 0)
    P = np.asfortranarray(np.zeros((8, 6, 0), np.float32))
    bp = mexct.CUDAmex_BP(P, 100.0, np.zeros(0, np.float32), geom)
    assert bp.shape == (8, 8, 6) and not bp.any()


@pytest.mark.parametrize("W,H", [(1, 1), (2, 1), (3, 5), (5, 2), (7, 64), (65, 3)])
@pytest.mark.parametrize("geom", [0, 1, 2])
def test_small_and_odd_sizes(gpu, W, H, geom):
    vol = _vol(W, H, W * 10 + H)
    ang = np.array([0.0, 0.4, 1.3, 2.9, 4.4], np.float32)
    fp = mexct.CUDAmex_FP(vol, 50.0, ang, ang.size, geom)
    assert np.array_equal(fp, O.fp(vol, 50.0, ang, ang.size, geom))
    P = np.asfortranarray(_rng(W + H).normal(size=(W, H, ang.size)).astype(np.float32))
    bp = mexct.CUDAmex_BP(P, 50.0, ang, geom)
    assert np.array_equal(bp, O.bp(P, 50.0, ang, geom))


@pytest.mark.parametrize("geom", [0, 1, 2])
def test_repeated_and_negative_angles(gpu, geom):
    W, H = 24, 10
    vol = _vol(W, H, 3)
    ang = np.array([-3.0, -0.5, -0.5, 0.0, 0.0, 7.5, 100.25], np.float32)
    fp = mexct.CUDAmex_FP(vol, 80.0, ang, ang.size, geom)
    assert np.array_equal(fp, O.fp(vol, 80.0, ang, ang.size, geom))
    assert np.array_equal(fp[:, :, 1], fp[:, :, 2]) and np.array_equal(fp[:, :, 3], fp[:, :, 4])
    bp = mexct.CUDAmex_BP(fp, 80.0, ang, geom)
    assert np.array_equal(bp, O.bp(fp, 80.0, ang, geom))


def test_rays_missing_the_volume(gpu):
    # SAD inside the volume's half-width: many cone rays leave the box or
    # walk backwards through clamped texels, exactly as the reference does
    W, H = 32, 16
    vol = _vol(W, H, 4)
    ang = np.linspace(0, 6, 7).astype(np.float32)
    for sad in (5.0, 12.0):
        for geom in (1, 2):
            fp = mexct.CUDAmex_FP(vol, sad, ang, ang.size, geom)
            assert np.array_equal(fp, O.fp(vol, sad, ang, ang.size, geom)), (sad, geom)


def test_nproj_below_angle_count(gpu):
    W, H = 16, 8
    vol = _vol(W, H, 5)
    ang = np.linspace(0, 3, 9).astype(np.float32)
    fp = mexct.CUDAmex_FP(vol, 200.0, ang, 4, 2)
    assert np.array_equal(fp, O.fp(vol, 200.0, ang[:4], 4, 2))


def test_tv_flat_volume_is_nan_like_reference(gpu):
    # a constant volume has a zero gradient, so c/|g| is inf and 0*inf NaN,
    # as in the reference (the TV step of OSC_TV_recon maps NaN to 0 after)
    x = np.asfortranarray(np.full((8, 8, 8), 0.5, np.float32))
    got = mexct.CUDAmex_TVmin_3D(x, 0.1, steps=1)
    ref = O.tvmin(x, 0.1, 1)
    assert np.array_equal(np.isnan(got), np.isnan(ref))


def test_gather_probe_unroll_variants(gpu):
    # the measured gather ceiling runs each probe unrolled 1, 2, 4 and 8 times:
    # every variant performs the same loads; other pattern bits are rejected
    from mexct import device

    ctx = device.Context(0)
    for storage in (mexct.F32, mexct.F16):
        for kind in (0, 1):
            counts = {ctx.probe_gather(storage, 40, 36, 6, kind | (k << 4)) for k in range(4)}
            assert len(counts) == 1 and counts.pop() > 0
    with pytest.raises(mexct.MexctError):
        ctx.probe_gather(mexct.F32, 40, 36, 6, 0x40)
    ctx.sync()
    ctx.close()


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16, mexct.BF16])
def test_loads_only_forms(gpu, storage):
    # The projectors' loads-only forms (their gather ceilings) run where the
    # z-fastest FP layouts and the cone BP apply, fold the same taps on
    # every call, reject the other cases, and leave the real operators'
    # results as they were.
    import torch

    from mexct import _lib as L
    from mexct import device

    W, H, n, sad = 48, 40, 12, 300.0
    ang = (np.arange(n) * 2 * np.pi / n).astype(np.float32)
    vol_np = _vol(W, H, 71)
    vol = torch.from_numpy(vol_np.T.copy()).cuda()
    ctx = device.Context(0)
    try:
        a = torch.empty((n, H, W), dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for geom in (0, 1, 2):
            ctx.fp_loads(vol, sad, ang, n, geom, a, storage=storage)
            ctx.fp_loads(vol, sad, ang, n, geom, b, storage=storage)
            torch.cuda.synchronize()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32))
        sino = torch.from_numpy(_rng(72).uniform(0, 1, (n, H, W)).astype(np.float32)).cuda()
        va = torch.empty((H, W, W), dtype=torch.float32, device="cuda")
        vb = torch.empty_like(va)
        ctx.bp_loads(sino, sad, ang, 2, va, storage=storage)
        ctx.bp_loads(sino, sad, ang, 2, vb, storage=storage)
        torch.cuda.synchronize()
        assert torch.equal(va.view(torch.int32), vb.view(torch.int32))
        with pytest.raises(mexct.MexctError) as e:
            ctx.bp_loads(sino, sad, ang, 0, va, storage=storage)
        assert e.value.code == L.ENOTSUP
        # the real operators afterwards: bit-identical to the oracle (f32)
        if storage == mexct.F32:
            ctx.fp(vol, sad, ang, n, 2, a)
            torch.cuda.synchronize()
            got = a.cpu().numpy().transpose(2, 1, 0)
            assert np.array_equal(got, O.fp(vol_np, sad, ang, n, 2))
    finally:
        ctx.close()
