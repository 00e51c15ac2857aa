"""Seeded synthetic inputs of the benchmark workloads (numpy, MATLAB layout).

All randomness comes from numpy's PCG64 generator with an explicit seed, so
the same seed gives the same arrays on every machine with this numpy.
"""
from __future__ import annotations

import numpy as np


def _axes(W, H):
    """Voxel-centre coordinates along x (= y) and z."""
    return (np.arange(W, dtype=np.float64) - (W - 1) / 2.0,
            np.arange(H, dtype=np.float64) - (H - 1) / 2.0)


def cylinder(W, H, radius, value=0.02):
    """Centred cylinder x^2 + y^2 <= r^2 on voxel centres, all slices
    (README.md:24-26 builds the same with meshgrid(-191.5:191.5))."""
    c = np.arange(W, dtype=np.float64) - (W - 1) / 2.0
    x, y = np.meshgrid(c, c, indexing="ij")
    m = (x * x + y * y) <= radius * radius
    vol = np.repeat((m * value).astype(np.float32)[:, :, None], H, axis=2)
    return np.asfortranarray(vol)


def ellipsoid_phantom(W, H, seed, n_ellipsoids, radius, base=0.02, axes=(5.0, 20.0),
                      values=(0.005, 0.03), centre_r=None, centre_z=None, additive=True):
    """Cylinder of `base` plus seeded random ellipsoids (z-rotated)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vol = cylinder(W, H, radius, base).astype(np.float64)
    c, zc = _axes(W, H)
    cr = radius * 0.8 if centre_r is None else centre_r
    cz = (H / 2.0) * 0.6 if centre_z is None else centre_z
    for _ in range(n_ellipsoids):
        a, b, cc = rng.uniform(axes[0], axes[1], size=3)
        val = rng.uniform(values[0], values[1])
        r = cr * np.sqrt(rng.uniform())
        phi = rng.uniform(0, 2 * np.pi)
        x0, y0, z0 = r * np.cos(phi), r * np.sin(phi), rng.uniform(-cz, cz)
        th = rng.uniform(0, np.pi)
        ct, st = np.cos(th), np.sin(th)
        # the z-rotated ellipsoid lies within |dx|, |dy| <= max(a, b), |dz| <= c
        rxy = max(a, b)
        ix = np.nonzero(np.abs(c - x0) <= rxy)[0]
        iy = np.nonzero(np.abs(c - y0) <= rxy)[0]
        iz = np.nonzero(np.abs(zc - z0) <= cc)[0]
        if ix.size == 0 or iy.size == 0 or iz.size == 0:
            continue
        sx = slice(ix[0], ix[-1] + 1)
        sy = slice(iy[0], iy[-1] + 1)
        sz = slice(iz[0], iz[-1] + 1)
        dx, dy, dz = np.meshgrid(c[sx] - x0, c[sy] - y0, zc[sz] - z0, indexing="ij")
        u = ct * dx + st * dy
        v = -st * dx + ct * dy
        m = (u / a) ** 2 + (v / b) ** 2 + (dz / cc) ** 2 <= 1.0
        box = vol[sx, sy, sz]
        if additive:
            box[m] += val
        else:
            box[m] = val
    return np.asfortranarray(vol.astype(np.float32))


def par_phantom_128(seed=1):
    """Parallel-beam workload volume: 128^3, cylinder r=48 of 0.02 plus 5
    ellipsoids (semi-axes 5-20, values 0.005-0.03)."""
    return ellipsoid_phantom(128, 128, seed, 5, 48.0, 0.02, (5.0, 20.0), (0.005, 0.03), 40.0, 40.0)


def cone_cylinder_384():
    """Cone-beam FDK workload volume: 384^3 cylinder r=140 of 0.02."""
    return cylinder(384, 384, 140.0, 0.02)


def osc_phantom(W=256, seed=3):
    """Few-view OSC-TV workload: 12 ellipsoids (0.01-0.03) in a r=110
    cylinder of 0.02 (W = 256 in the full-size case)."""
    s = W / 256.0
    return ellipsoid_phantom(W, W, seed, 12, 110.0 * s, 0.02, (8.0 * s, 30.0 * s), (0.01, 0.03),
                             80.0 * s, 0.3 * W, additive=False)


def poisson_transmission(fp, I0=1e5, seed=33):
    """pre = Poisson(I0), post = Poisson(I0 * exp(-fp)), as float32 arrays."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = I0 * np.exp(-np.asarray(fp, dtype=np.float64))
    pre = rng.poisson(I0, size=fp.shape).astype(np.float32)
    post = rng.poisson(lam).astype(np.float32)
    return np.asfortranarray(pre), np.asfortranarray(post)


def angles_uniform(n, span):
    """k*span/n, k < n, as float32."""
    return (np.arange(n, dtype=np.float64) * span / n).astype(np.float32)


def angles_linspace(n, end):
    """single(linspace(0, end, n)) (README.md:30): endpoints included."""
    return np.linspace(0.0, end, n).astype(np.float32)


def parallel_pts(W, H, dist=1000.0):
    """general3D pts reproducing parallel-beam rays (general3D_projection_test.m:13-29)."""
    x1 = np.full((W, H), -dist, dtype=np.float32)
    y1 = np.repeat((np.arange(W) - (W - 1) / 2.0)[:, None], H, axis=1).astype(np.float32)
    z1 = np.repeat((np.arange(H) - (H - 1) / 2.0)[None, :], W, axis=0).astype(np.float32)
    pts = np.stack([x1, y1, z1, -x1, y1, z1], axis=2)
    return np.asfortranarray(pts.astype(np.float32))


def helix_pts(W_det, H_det, n_views, vol_side, radius=600.0, turns=2.0, pitch_per_turn=None,
              pixel=None, jitter=0.25, seed=4):
    """Per-view general-3D rays of a helical cone scan: the source at
    (R cos t, R sin t, z(t)) with z rising one volume height per turn, a flat
    W_det x H_det detector centred opposite the source at radius R, each
    pixel jittered by U[-jitter, jitter] pixel.  Returns [W, H, 6, nViews]
    with angles 0 (the rays are already placed per view)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pitch = float(vol_side) if pitch_per_turn is None else pitch_per_turn
    pix = (2.5 * 256.0 / W_det) if pixel is None else pixel
    out = np.zeros((W_det, H_det, 6, n_views), dtype=np.float32, order="F")
    iu = np.arange(W_det) - (W_det - 1) / 2.0
    iv = np.arange(H_det) - (H_det - 1) / 2.0
    for k in range(n_views):
        t = k * 2.0 * np.pi * turns / n_views
        zs = -vol_side / 2.0 + pitch * t / (2.0 * np.pi)
        zs = ((zs + vol_side / 2.0) % vol_side) - vol_side / 2.0
        c, s = np.cos(t), np.sin(t)
        src = np.array([radius * c, radius * s, zs])
        du = iu[:, None] + rng.uniform(-jitter, jitter, size=(W_det, H_det))
        dv = iv[None, :] + rng.uniform(-jitter, jitter, size=(W_det, H_det))
        # detector plane at distance R on the far side, u along (-s, c), v along z
        dx = -radius * c + (-s) * du * pix
        dy = -radius * s + c * du * pix
        dz = zs + dv * pix
        out[:, :, 0, k] = src[0]
        out[:, :, 1, k] = src[1]
        out[:, :, 2, k] = src[2]
        out[:, :, 3, k] = dx
        out[:, :, 4, k] = dy
        out[:, :, 5, k] = dz
    return out
