"""The forward projector's address paths past the default layouts.

The z-fastest projector reads its x-pair (f32) or corner-quad (2-byte)
volume copy through 32-bit buffer offsets while the copy stays below 4 GiB
and through 64-bit global addresses past that (about 810^3 f32, 1024^3
included).  A detector with (W+4)(H+4) >= 2^23 goes to the x-fastest
projector.  These tests reach every one of those paths:

- MEXCT_FP_LAYOUT forces the 64-bit reads, or the one-value-per-element
  copy, on small volumes, whose results must equal the default layout's bit
  for bit (FP, batched FP, general-3D FP and the OSC projection);
- a 1024^3 f32 cone FP, past the 4 GiB copy, is checked on sampled rays
  against the oracle bit for bit, and its f16 form against the f32 result;
- a 64 x 64 x 123400 volume goes to the x-fastest projector, checked on
  sampled rays against the oracle bit for bit (the cone rays of the rows
  the reference computes cross the volume near slice 30850);
- past 2^31 haloed elements, up to the reference's unsigned 32-bit index
  range: a 1296^3 FP (1300^3 haloed) on sampled rays, and a 512^2 x 8200
  view BP (the reference walks any number of views) on sampled voxel
  columns, each against the oracle bit for bit and in f16 against f32.
"""
import sys

import numpy as np
import pytest

import mexct
import oracle as O
from conftest import rel_rmse
from mexct import device

pytestmark = pytest.mark.gpu

LAYOUTS = ["wide64", "plain"]
STORAGES = [mexct.F32, mexct.F16, mexct.BF16]


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda(gpu):
    import torch
    return torch


def _check(got, ref, layout, storage, what=""):
    """The 64-bit reads must give the default layout's bits.  The plain copy
    gives them for f32; with 2-byte storage its taps are widened to f32
    before the first lerp, where the corner-quad path lerps in packed f16,
    so it agrees to the f16 storage tolerance."""
    if layout == "wide64" or storage == mexct.F32:
        assert torch_equal(got, ref), what
    else:
        g = got.double().cpu().numpy()
        r = ref.double().cpu().numpy()
        m = float(np.abs(g - r).max() / np.abs(r).max())
        e = rel_rmse(g, r)
        _say(f"  {what} plain vs default: rel max {m:.2e} relRMSE {e:.2e}")
        assert m <= 1e-3 and e <= 2e-4, (what, m, e)


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))


def _forced(monkeypatch, layout, fn):
    """fn() with MEXCT_FP_LAYOUT=layout, the variable removed afterwards."""
    monkeypatch.setenv("MEXCT_FP_LAYOUT", layout)
    try:
        return fn()
    finally:
        monkeypatch.delenv("MEXCT_FP_LAYOUT")


@pytest.mark.parametrize("geom,sad", [(0, 500.0), (1, 150.0), (2, 300.0)])
@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("layout", LAYOUTS)
def test_forced_layout_fp_equals_default(torch_cuda, monkeypatch, layout, storage, geom, sad):
    torch = torch_cuda
    W, H, n = 72, 40, 24
    _say(f"fp layout={layout} storage={storage} geom={geom} W={W} H={H} n={n}")
    ang = (np.arange(n) * (2 * np.pi / n) + 0.05).astype(np.float32)
    vol = torch.from_numpy(_rng(W + geom).uniform(0, 1, (H, W, W)).astype(np.float32)).cuda()
    ctx = device.Context(0)
    ref = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp(vol, sad, ang, n, geom, ref, storage=storage)
    got = torch.full((n, H, W), -1.0, dtype=torch.float32, device="cuda")
    _forced(monkeypatch, layout, lambda: ctx.fp(vol, sad, ang, n, geom, got, storage=storage))
    torch.cuda.synchronize()
    ctx.close()
    _check(got, ref, layout, storage, "fp")


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
def test_forced_wide64_batch_equals_default(torch_cuda, monkeypatch, storage):
    torch = torch_cuda
    G, W, H, n, geom, sad = 3, 56, 33, 20, 2, 150.0
    _say(f"fp batch wide64 storage={storage} G={G} W={W} H={H} n={n}")
    ang = (np.arange(n) * (2 * np.pi / n) + 0.05).astype(np.float32)
    vols = torch.from_numpy(_rng(7).uniform(0, 1, (G, H, W, W)).astype(np.float32)).cuda()
    ctx = device.Context(0)
    ref = torch.zeros((G, n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp_batch(vols, sad, ang, n, geom, ref, storage=storage)
    got = torch.full((G, n, H, W), -1.0, dtype=torch.float32, device="cuda")
    _forced(monkeypatch, "wide64",
            lambda: ctx.fp_batch(vols, sad, ang, n, geom, got, storage=storage))
    torch.cuda.synchronize()
    ctx.close()
    assert torch.equal(got, ref)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("layout", LAYOUTS)
def test_forced_layout_g3d_fp_equals_default(torch_cuda, monkeypatch, layout, storage):
    torch = torch_cuda
    side, nsl, Wd, Hd, n = 40, 36, 24, 20, 4
    _say(f"g3d fp layout={layout} storage={storage} side={side} nsl={nsl}")
    r = _rng(11)
    vol = torch.from_numpy(r.uniform(0, 1, (nsl, side, side)).astype(np.float32)).cuda()
    pts = torch.from_numpy(r.uniform(-45, 45, (n, 6, Hd, Wd)).astype(np.float32)).cuda()
    ang = (np.arange(n) * 0.6).astype(np.float32)
    ctx = device.Context(0)
    ref = torch.zeros((n, Hd, Wd), dtype=torch.float32, device="cuda")
    ctx.g3d_fp(vol, ang, pts, n, ref, storage=storage)
    got = torch.full((n, Hd, Wd), -1.0, dtype=torch.float32, device="cuda")
    _forced(monkeypatch, layout, lambda: ctx.g3d_fp(vol, ang, pts, n, got, storage=storage))
    torch.cuda.synchronize()
    ctx.close()
    _check(got, ref, layout, storage, "g3d fp")


@pytest.mark.parametrize("geom,sad", [(0, 500.0), (2, 200.0)])
@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("layout", LAYOUTS)
def test_forced_layout_osc_equals_default(torch_cuda, monkeypatch, layout, storage, geom, sad):
    torch = torch_cuda
    W, H, n, S = 48, 20, 24, 4
    _say(f"osc layout={layout} storage={storage} geom={geom} W={W} H={H} n={n}")
    r = _rng(21 + geom)
    mu0 = torch.from_numpy(r.uniform(0.0, 0.02, (H, W, W)).astype(np.float32)).cuda()
    pre0 = torch.from_numpy(r.uniform(0.5, 1.0, (n, H, W)).astype(np.float32)).cuda()
    post0 = torch.from_numpy(r.uniform(0.0, 0.1, (n, H, W)).astype(np.float32)).cuda()
    ang = (np.arange(n) * (2 * np.pi / n)).astype(np.float32)
    picks = np.array([0, 2, 1, 3], dtype=np.int32)
    ctx = device.Context(0)
    a = [mu0.clone(), pre0.clone(), post0.clone()]
    ctx.osc_iter(a[0], a[1], a[2], S, picks, ang, sad, geom, storage=storage)
    b = [mu0.clone(), pre0.clone(), post0.clone()]
    _forced(monkeypatch, layout,
            lambda: ctx.osc_iter(b[0], b[1], b[2], S, picks, ang, sad, geom, storage=storage))
    torch.cuda.synchronize()
    ctx.close()
    for x, y, name in zip(a, b, ("mu", "pre", "post")):
        _check(y, x, layout, storage, "osc " + name)


def _copy_bytes_f32(W, H):
    """Bytes of the x-pair copy of a W x W x H f32 volume."""
    return (W + 3) * (W + 4) * (H + 4) * 2 * 4


def test_fp_1024_cube_past_4gib_sampled(torch_cuda):
    torch = torch_cuda
    W = H = 1024
    n, geom, sad = 4, 2, 2000.0
    assert _copy_bytes_f32(W, H) >= 1 << 32  # the 64-bit read path
    _say(f"fp 1024^3 f32 cone n={n}: 64-bit reads of the x-pair copy")
    g = torch.Generator(device="cuda")
    g.manual_seed(1024)
    vol = torch.rand((H, W, W), generator=g, device="cuda", dtype=torch.float32)
    ang = (np.arange(n) * (2 * np.pi / n) + 0.1).astype(np.float32)
    ctx = device.Context(0)
    out = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp(vol, sad, ang, n, geom, out)
    torch.cuda.synchronize()
    fp = out.cpu().numpy()  # (view, v, u)
    _say("fp 1024^3 f16 storage: 64-bit reads of the corner-quad copy")
    out16 = torch.zeros_like(out)
    ctx.fp(vol, sad, ang, n, geom, out16, storage=mexct.F16)
    torch.cuda.synchronize()
    fp16 = out16.cpu().numpy()
    ctx.close()
    vnp = vol.cpu().numpy()
    del vol, out, out16
    torch.cuda.empty_cache()
    rng = _rng(99)
    iu, iv, ik = rng.integers(0, W, 4096), rng.integers(0, H, 4096), rng.integers(0, n, 4096)
    # the oracle takes the MATLAB layout vol[x, y, z]: the (z, y, x) array transposed
    ref = O.fp_rays(vnp.T, sad, ang, geom, iu, iv, ik)
    assert np.array_equal(fp[ik, iv, iu], ref)
    assert fp[ik, iv, iu].max() > 100.0  # the rays cross the volume
    e = rel_rmse(fp16, fp)
    _say(f"f16 storage relRMSE vs f32: {e:.3e}")
    assert e <= 1e-3


@pytest.mark.parametrize("geom,sad", [(0, 1000.0), (1, 300.0), (2, 1.0e5)])
def test_fp_tall_volume_xfastest_sampled(torch_cuda, geom, sad):
    torch = torch_cuda
    W, H, n = 64, 123400, 2
    assert (W + 4) * (H + 4) >= 1 << 23  # past the z-fastest index range
    _say(f"fp {W}x{W}x{H} geom={geom} n={n}: x-fastest projector")
    g = torch.Generator(device="cuda")
    g.manual_seed(64 + geom)
    vol = torch.rand((H, W, W), generator=g, device="cuda", dtype=torch.float32)
    ang = np.array([0.3, 1.9], dtype=np.float32)
    ctx = device.Context(0)
    out = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp(vol, sad, ang, n, geom, out)
    torch.cuda.synchronize()
    ctx.close()
    fp = out.cpu().numpy()
    vnp = vol.cpu().numpy()
    del vol, out
    # the reference's grid covers detector rows below 32 * ceil(W / 32) only
    # (CUDAmex_FP.cu:434-440); the rows above stay zero
    rows = min(H, 32 * ((W + 31) // 32))
    assert not fp[:, rows:, :].any()
    rng = _rng(5 + geom)
    iu, iv, ik = rng.integers(0, W, 4096), rng.integers(0, rows, 4096), rng.integers(0, n, 4096)
    ref = O.fp_rays(vnp.T, sad, ang, geom, iu, iv, ik)
    assert np.array_equal(fp[ik, iv, iu], ref)
    assert fp[ik, iv, iu].max() > 10.0


def test_fp_past_2_31_haloed_sampled(torch_cuda):
    """1296^3: (W+4)^2 (H+4) = 1300^3 haloed elements, past 2^31.  The
    reference indexes its volume in unsigned 32 bits (CUDAmex_FP.cu:394), so
    it takes volumes up to 2^32 elements; the x-pair and corner-quad copies
    read through 64-bit addresses from 32-bit element indices."""
    torch = torch_cuda
    W = H = 1296
    n, geom, sad = 4, 2, 2600.0
    assert (W + 4) ** 2 * (H + 4) >= 1 << 31
    _say(f"fp {W}^3 f32 cone n={n}: past 2^31 haloed elements")
    g = torch.Generator(device="cuda")
    g.manual_seed(1296)
    vol = torch.rand((H, W, W), generator=g, device="cuda", dtype=torch.float32)
    ang = (np.arange(n) * (2 * np.pi / n) + 0.2).astype(np.float32)
    ctx = device.Context(0)
    out = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    ctx.fp(vol, sad, ang, n, geom, out)
    torch.cuda.synchronize()
    fp = out.cpu().numpy()
    _say(f"fp {W}^3 f16 storage")
    out16 = torch.zeros_like(out)
    ctx.fp(vol, sad, ang, n, geom, out16, storage=mexct.F16)
    torch.cuda.synchronize()
    fp16 = out16.cpu().numpy()
    ctx.close()
    vnp = vol.cpu().numpy()
    del vol, out, out16
    torch.cuda.empty_cache()
    rng = _rng(1300)
    iu, iv, ik = rng.integers(0, W, 4096), rng.integers(0, H, 4096), rng.integers(0, n, 4096)
    ref = O.fp_rays(vnp.T, sad, ang, geom, iu, iv, ik)
    assert np.array_equal(fp[ik, iv, iu], ref)
    assert fp[ik, iv, iu].max() > 100.0
    e = rel_rmse(fp16, fp)
    _say(f"f16 storage relRMSE vs f32: {e:.3e}")
    assert e <= 1e-3


def test_bp_past_2_31_haloed_sinogram_sampled(torch_cuda):
    """512^2 detector x 8200 views: (W+4)(H+4)(n+4) past 2^31 haloed
    elements.  The reference walks any number of views, one projection at a
    time (CUDAmex_BP.cu:287-291); here the views run in launches below 2 GiB
    of sinogram copy that continue each other's sums."""
    torch = torch_cuda
    W = H = 512
    n, geom, sad = 8200, 2, 1500.0
    assert (W + 4) * (H + 4) * (n + 4) >= 1 << 31
    _say(f"bp {W}^2 x {n} views f32 cone: past 2^31 haloed sinogram elements")
    g = torch.Generator(device="cuda")
    g.manual_seed(8200)
    sino = torch.rand((n, H, W), generator=g, device="cuda", dtype=torch.float32)
    ang = (np.arange(n) * (2 * np.pi / n)).astype(np.float32)
    ctx = device.Context(0)
    out = torch.zeros((H, W, W), dtype=torch.float32, device="cuda")
    ctx.bp(sino, sad, ang, geom, out)
    torch.cuda.synchronize()
    rec = out.cpu().numpy()  # [z, y, x]
    _say("bp f16 storage")
    out16 = torch.zeros_like(out)
    ctx.bp(sino, sad, ang, geom, out16, storage=mexct.F16)
    torch.cuda.synchronize()
    rec16 = out16.cpu().numpy()
    ctx.close()
    P = sino.cpu().numpy()
    del sino, out, out16
    torch.cuda.empty_cache()
    rng = _rng(516)
    hw = W // 2 - 1
    cols = []
    while len(cols) < 16:
        a, b = rng.integers(0, W, 2)
        if (a - hw) ** 2 + (b - hw) ** 2 < hw * hw:
            cols.append((a, b))
    cx = np.array([c[0] for c in cols])
    cy = np.array([c[1] for c in cols])
    # the oracle takes the MATLAB layout sino[u, v, view]: the (view, v, u) array transposed
    ref = O.bp_cols(P.T, sad, ang, geom, cx, cy)
    assert np.array_equal(rec[:, cy, cx].T, ref)
    assert ref.max() > 100.0
    e = rel_rmse(rec16, rec)
    _say(f"f16 storage relRMSE vs f32: {e:.3e}")
    assert e <= 1e-3
