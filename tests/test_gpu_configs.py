"""The workloads at their real sizes.

- parallel-beam FP then BP of the 128^3 phantom, 180 views, checked in full;
- cone-beam FDK on the 384^3 cylinder, 512 views: FP, filter and BP at full
  size, checked on sampled rays, views and voxel columns (the full oracle
  would take minutes);
- helical general-3D FP/BP on a 320^3 volume, 360 views of a 256x256
  jittered detector: the full FP runs on the device, three views are
  checked against the oracle; the BP of those views within the atomics
  tolerance.
The few-view OSC-TV case (256^3, 60 views) is checked at 64^3 against the
FP64 oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

import mexct
import oracle as O
from conftest import rel_errors
from mexct import phantoms

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def test_parallel_128_fp_bp_full(gpu):
    vol = phantoms.par_phantom_128(seed=1)
    ang = phantoms.angles_uniform(180, np.pi)
    fp = mexct.CUDAmex_FP(vol, 1000.0, ang, 180, 0)
    assert np.array_equal(fp, O.fp(vol, 1000.0, ang, 180, 0))
    bp = mexct.CUDAmex_BP(fp, 1000.0, ang, 0)
    assert np.array_equal(bp, O.bp(fp, 1000.0, ang, 0))


def test_cone_384_fdk_sampled(gpu):
    W = H = 384
    n = 512
    vol = phantoms.cone_cylinder_384()
    ang = phantoms.angles_linspace(n, 2 * np.pi)
    fp = mexct.CUDAmex_FP(vol, 1000.0, ang, n, 2)
    rng = _rng(12)
    ix, iy, iv = rng.integers(0, W, 4096), rng.integers(0, H, 4096), rng.integers(0, n, 4096)
    assert np.array_equal(fp[ix, iy, iv], O.fp_rays(vol, 1000.0, ang, 2, ix, iy, iv))
    geo = {"type": "cone3d", "SAD": 1000.0}
    filt = mexct.preFilterSinogram(fp, geo, ang, "hamming", 1.0)
    # the filter of views 0..7 depends only on them (angle step from the
    # previous view, the first view's from the next)
    ref = O.prefilter(np.asfortranarray(fp[:, :, :8]), "cone3d", 1000.0, ang[:8], "hamming", 1.0)
    m, l2 = rel_errors(filt[:, :, :8], ref)
    assert m <= 1e-6 and l2 <= 1e-7, (m, l2)
    rec = mexct.CUDAmex_BP(filt, 1000.0, ang, 2)
    hw = W // 2 - 1
    cols = []
    while len(cols) < 32:
        a, b = rng.integers(0, W, 2)
        if (a - hw) ** 2 + (b - hw) ** 2 < hw * hw:
            cols.append((a, b))
    cx = np.array([c[0] for c in cols])
    cy = np.array([c[1] for c in cols])
    assert np.array_equal(rec[cx, cy, :], O.bp_cols(filt, 1000.0, ang, 2, cx, cy))
    assert not rec[0, 0, :].any()  # outside the circle


def test_helix_g3d_320(gpu):
    side, nv, Wd = 320, 360, 256
    vol = phantoms.ellipsoid_phantom(side, side, 4, 12, 140.0, 0.02, (8.0, 40.0), (0.01, 0.03),
                                     100.0, 120.0)
    pts = phantoms.helix_pts(Wd, Wd, nv, side)
    ang = np.zeros(nv, np.float32)
    fp = mexct.CUDAmex_general3D_FP(vol, ang, pts, nv)
    assert fp.shape == (Wd, Wd, nv) and np.isfinite(fp).all() and fp.any()
    sub = [0, 137, 359]
    psub = np.asfortranarray(pts[:, :, :, sub])
    ref = O.g3d_fp(vol, ang[:3], psub, 3)
    assert np.array_equal(fp[:, :, sub], ref)
    P = np.asfortranarray(fp[:, :, sub])
    got = mexct.CUDAmex_general3D_BP(P, ang[:3], psub, side, side)
    refb = O.g3d_bp(P, ang[:3], psub, side, side)
    m, l2 = rel_errors(got, refb)
    print(f"helix g3D BP (3 views): rel_max={m:.3e} rel_l2={l2:.3e}")
    assert m <= 1e-5 and l2 <= 1e-6


def test_bp_views_in_chunks(gpu):
    # 512^2 detector x 1100 views: the row-paired sinogram (2.34 GB) passes
    # 32-bit buffer offsets, so the views run in two chunks that continue
    # each other's sums
    W = H = 512
    n = 1100
    rng = _rng(21)
    P = np.asfortranarray(rng.uniform(0, 1, (W, H, n)).astype(np.float32))
    ang = phantoms.angles_uniform(n, 2 * np.pi)
    rec = mexct.CUDAmex_BP(P, 1000.0, ang, 2)
    hw = W // 2 - 1
    cols = []
    while len(cols) < 12:
        a, b = rng.integers(0, W, 2)
        if (a - hw) ** 2 + (b - hw) ** 2 < hw * hw:
            cols.append((a, b))
    cx = np.array([c[0] for c in cols])
    cy = np.array([c[1] for c in cols])
    assert np.array_equal(rec[cx, cy, :], O.bp_cols(P, 1000.0, ang, 2, cx, cy))
