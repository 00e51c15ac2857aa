"""Pins of the CPU oracle: closed forms, invariants and independent
computations that a plausible transcription mistake (dropped term, wrong
sign or index, transposed operand) would fail.  The reference ships no
tests or golden vectors, so these are the oracle's anchors.  CPU only."""
import numpy as np
import pytest

import oracle as O


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


# ------------------------------------------------------------------ sampler
def _numpy_trilinear(T, x, y, z, mode):
    """Independent float64 trilinear fetch with the public texture rule."""
    out = 0.0
    xb, yb, zb = x - 0.5, y - 0.5, z - 0.5
    i, j, k = int(np.floor(xb)), int(np.floor(yb)), int(np.floor(zb))
    a, b, c = xb - i, yb - j, zb - k
    for di, wi in ((0, 1 - a), (1, a)):
        for dj, wj in ((0, 1 - b), (1, b)):
            for dk, wk in ((0, 1 - c), (1, c)):
                ii, jj, kk = i + di, j + dj, k + dk
                if mode == O.CLAMP:
                    ii = min(max(ii, 0), T.shape[0] - 1)
                    jj = min(max(jj, 0), T.shape[1] - 1)
                    kk = min(max(kk, 0), T.shape[2] - 1)
                    v = T[ii, jj, kk]
                elif 0 <= ii < T.shape[0] and 0 <= jj < T.shape[1] and 0 <= kk < T.shape[2]:
                    v = T[ii, jj, kk]
                else:
                    v = 0.0
                out += wi * wj * wk * float(v)
    return out


def test_tex3_texel_centres_exact():
    T = f32(rng(1).uniform(size=(5, 4, 3)))
    for mode in (O.CLAMP, O.BORDER):
        for i in range(5):
            for j in range(4):
                for k in range(3):
                    assert O.tex3(T, i + 0.5, j + 0.5, k + 0.5, mode) == T[i, j, k]


def test_tex3_matches_independent_trilinear():
    T = f32(rng(2).uniform(-1, 1, size=(6, 5, 4)))
    r = rng(3)
    for mode in (O.CLAMP, O.BORDER):
        for _ in range(400):
            x, y, z = r.uniform(-2, 8), r.uniform(-2, 7), r.uniform(-2, 6)
            x, y, z = float(np.float32(x)), float(np.float32(y)), float(np.float32(z))
            assert O.tex3(T, x, y, z, mode) == pytest.approx(_numpy_trilinear(T, x, y, z, mode),
                                                              rel=1e-5, abs=1e-6)


def test_tex3_address_modes_outside():
    T = f32(rng(4).uniform(0.5, 1.0, size=(4, 4, 4)))
    assert O.tex3(T, -7.0, 1.5, 2.5, O.CLAMP) == T[0, 1, 2]
    assert O.tex3(T, 13.0, 1.5, 2.5, O.CLAMP) == T[3, 1, 2]
    assert O.tex3(T, -7.0, 1.5, 2.5, O.BORDER) == 0.0
    # half a texel outside: border blends with 0, clamp does not
    assert O.tex3(T, 0.0, 1.5, 2.5, O.BORDER) == pytest.approx(0.5 * T[0, 1, 2], rel=1e-7)
    assert O.tex3(T, 0.0, 1.5, 2.5, O.CLAMP) == T[0, 1, 2]


def test_texture_weight_bits_mode():
    # 8 fraction bits: the filter fraction rounds to the nearest multiple of
    # 1/256; texel centres (fraction 0) and the exact mode are unchanged
    T = f32(np.arange(10, dtype=np.float32).reshape(10, 1))
    V = f32(rng(7).uniform(size=(6, 5, 4)))
    for x in (2.3, 4.5017, 7.998, 1.0):
        frac = (x - 0.5) - np.floor(x - 0.5)
        want = np.floor(x - 0.5) + np.round(frac * 256) / 256
        with O.texture_weight_bits(8):
            assert O.tex2(T, x, 0.5) == pytest.approx(want, abs=1e-6)
        assert O.tex2(T, x, 0.5) == pytest.approx(x - 0.5, abs=1e-6)
    with O.texture_weight_bits(8):
        q = O.tex3(V, 2.3, 1.7, 2.9, O.CLAMP)
        c = O.tex3(V, 2.5, 1.5, 3.5, O.CLAMP)
    assert O.tex3(V, 2.5, 1.5, 3.5, O.CLAMP) == c == V[2, 1, 3]
    assert q == pytest.approx(_numpy_trilinear(V, 2.3 - 0.3 + 77 / 256, 1.7 - 0.2 + 51 / 256,
                                               2.9 - 0.4 + 102 / 256, O.CLAMP), abs=1e-6)
    assert O.threads() > 0 and O.lib().orc_get_texture_weight_bits() == 0


def test_texture_weight_gap_small_parallel():
    # exact against 8-bit weights on a small parallel FP/BP: the gap is far
    # below the FP16 storage error and within the FP32 parity bar's order
    import mexct.phantoms as ph

    vol = f32(ph.ellipsoid_phantom(32, 32, 1, 3, 12.0))
    ang = ph.angles_uniform(24, np.pi)
    a = O.fp(vol, 1000.0, ang, 24, 0)
    with O.texture_weight_bits(8):
        b = O.fp(vol, 1000.0, ang, 24, 0)
        bb = O.bp(a, 1000.0, ang, 0)
    ba = O.bp(a, 1000.0, ang, 0)
    for x, y in ((b, a), (bb, ba)):
        d = np.abs(x.astype(np.float64) - y)
        assert 0 < d.max() / np.abs(y).max() < 1e-3


def test_tex2_linear_function_reproduced():
    a, b, c = 0.75, -1.25, 3.0
    u, v = np.meshgrid(np.arange(9), np.arange(7), indexing="ij")
    T = f32(a * u + b * v + c)
    r = rng(5)
    for _ in range(200):
        x, y = float(np.float32(r.uniform(0.5, 8.5))), float(np.float32(r.uniform(0.5, 6.5)))
        assert O.tex2(T, x, y) == pytest.approx(a * (x - 0.5) + b * (y - 0.5) + c, abs=2e-5)


# ------------------------------------------------------------------ FP
def test_fp_parallel_beta0_is_column_sum():
    # beta = 0: rays run along -x through texel centres in y and z, 0.5 steps;
    # with zero x-faces the ray sum is sum_x vol[x, u, v]
    W, H = 24, 10
    vol = np.zeros((W, W, H), np.float32, order="F")
    vol[1:-1] = rng(6).uniform(size=(W - 2, W, H))
    p = O.fp(vol, 1000.0, np.zeros(1, np.float32), 1, 0)[:, :, 0]
    np.testing.assert_allclose(p, vol.astype(np.float64).sum(axis=0), rtol=2e-5)


def test_fp_parallel_beta0_edge_term():
    # non-zero faces add (T[W-1] - T[0]) / 4 when 2W samples are taken
    W, H = 16, 4
    vol = f32(rng(7).uniform(size=(W, W, H)))
    p, n = O.fp(vol, 1000.0, np.zeros(1, np.float32), 1, 0, return_samples=True)
    S = vol.astype(np.float64).sum(axis=0)
    edge = (vol[-1].astype(np.float64) - vol[0]) / 4.0
    assert n == 2 * W * W * H
    np.testing.assert_allclose(p[:, :, 0], S + edge, rtol=1e-4)


def test_fp_parallel_cylinder_chord():
    W, H, R, mu = 64, 4, 20.0, 0.02
    c = np.arange(W) - (W - 1) / 2.0
    x, y = np.meshgrid(c, c, indexing="ij")
    vol = f32(np.repeat(((x * x + y * y) <= R * R)[:, :, None] * mu, H, axis=2))
    ang = np.array([0.0, 0.7, 1.9], np.float32)
    p = O.fp(vol, 1000.0, ang, 3, 0)
    chord = 2 * mu * np.sqrt(np.clip(R * R - c * c, 0, None))
    for k in range(3):
        assert np.abs(p[:, 1, k] - chord).max() < 1.5 * mu * 2
        assert abs(p[:, 1, k].sum() - chord.sum()) < 0.01 * chord.sum()


def test_fp_linearity_and_zero():
    W, H = 20, 12
    a, b = f32(rng(8).uniform(size=(W, W, H))), f32(rng(9).uniform(size=(W, W, H)))
    ang = np.array([0.3, 1.1, 2.5], np.float32)
    for geom in (0, 1, 2):
        pa = O.fp(a, 200.0, ang, 3, geom, f64=True)
        pb = O.fp(b, 200.0, ang, 3, geom, f64=True)
        pab = O.fp(f32(2 * a.astype(np.float64) + 3 * b), 200.0, ang, 3, geom, f64=True)
        np.testing.assert_allclose(pab, 2 * pa + 3 * pb, rtol=1e-5, atol=1e-6)
        assert not O.fp(np.zeros((W, W, H), np.float32, order="F"), 200.0, ang, 3, geom).any()


def test_fp_parallel_opposite_views_mirror():
    # beta and beta + pi see the same line integrals with u mirrored; the
    # sample offsets along the ray differ, so the content is kept smooth
    W, H = 32, 6
    c = np.arange(W) - (W - 1) / 2.0
    x, y = np.meshgrid(c, c, indexing="ij")
    blob = np.exp(-((x - 3) ** 2 + (y + 2) ** 2) / (2 * 4.0 ** 2))
    vol = f32(np.repeat(blob[:, :, None], H, axis=2))
    p = O.fp(vol, 1000.0, np.array([0.4, 0.4 + np.pi], np.float32), 2, 0, f64=True)
    np.testing.assert_allclose(p[:, :, 1], p[::-1, :, 0], rtol=0, atol=2e-3 * p.max())


def test_fp_f32_close_to_f64_twin():
    W, H = 24, 16
    vol = f32(rng(11).uniform(size=(W, W, H)))
    ang = np.array([0.2, 1.0, 2.2, 4.0], np.float32)
    for geom in (0, 1, 2):
        a = O.fp(vol, 150.0, ang, 4, geom)
        b = O.fp(vol, 150.0, ang, 4, geom, f64=True)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4


# ------------------------------------------------------------------ BP
def _circle(W):
    hw = W // 2 - 1
    ix, iy = np.meshgrid(np.arange(W), np.arange(W), indexing="ij")
    return (ix - hw) ** 2 + (iy - hw) ** 2 < hw * hw


def test_bp_parallel_single_view_closed_form():
    # beta = 0: voxel (ix,iy,iz) samples texel (u=iy, v=iz) exactly
    W, H = 18, 7
    P = f32(rng(12).normal(size=(W, H, 1)))
    out = O.bp(P, 1000.0, np.zeros(1, np.float32), 0)
    m = _circle(W)
    expect = np.where(m[:, :, None], P[None, :, :, 0], 0.0)
    np.testing.assert_array_equal(out, expect.astype(np.float32))


def test_bp_constant_sinogram():
    W, H, n, c, SR = 20, 9, 7, 0.75, 150.0
    P = np.full((W, H, n), c, np.float32, order="F")
    ang = (np.arange(n) * 0.7).astype(np.float32)
    m = _circle(W)
    out = O.bp(P, SR, ang, 0)
    np.testing.assert_allclose(out[m], c * n, rtol=1e-6)
    assert not out[~m].any()
    # cone: every tap is c (clamp), so out = c * sum_views U^2
    out = O.bp(P, SR, ang, 2)
    hw = W // 2 - 1
    ix, iy = np.nonzero(m)
    x = ix - (hw + 0.5)
    y = iy - (hw + 0.5)
    s_, c_ = np.sin(ang.astype(np.float32)), np.cos(ang.astype(np.float32))
    U = SR / (SR - (x[:, None] * c_[None] + y[:, None] * s_[None]))
    np.testing.assert_allclose(out[ix, iy, 3], c * (U * U).sum(axis=1), rtol=2e-6)


def test_bp_f32_close_to_f64_twin():
    W, H, n = 22, 14, 6
    P = f32(rng(13).normal(size=(W, H, n)))
    ang = (np.arange(n) * 1.1).astype(np.float32)
    for geom in (0, 1, 2):
        a = O.bp(P, 120.0, ang, geom)
        b = O.bp(P, 120.0, ang, geom, f64=True)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5


def test_bp_cols_match_full():
    W, H, n = 20, 10, 5
    P = f32(rng(14).normal(size=(W, H, n)))
    ang = (np.arange(n) * 0.9).astype(np.float32)
    full = O.bp(P, 100.0, ang, 2)
    ix = np.array([3, 10, 9, 15])
    iy = np.array([9, 10, 2, 12])
    np.testing.assert_array_equal(O.bp_cols(P, 100.0, ang, 2, ix, iy), full[ix, iy, :])
    vol = f32(rng(15).uniform(size=(W, W, H)))
    fp = O.fp(vol, 100.0, ang, n, 2)
    r = O.fp_rays(vol, 100.0, ang, 2, ix, iy % H, np.arange(4) % n)
    np.testing.assert_array_equal(r, fp[ix, iy % H, np.arange(4) % n])


# ------------------------------------------------------------------ general 3D
def test_g3d_fp_parallel_pts_equals_parallel_fp():
    # general3D_projection_test.m:13-29 builds parallel rays as pts
    import mexct.phantoms as ph

    W, H = 24, 12
    vol = np.zeros((W, W, H), np.float32, order="F")
    vol[3:-3, 3:-3, 3:-3] = rng(16).uniform(size=(W - 6, W - 6, H - 6))
    ang = (np.arange(4) * np.pi / 2).astype(np.float32)
    a = O.g3d_fp(vol, ang, ph.parallel_pts(W, H), 4)
    b = O.fp(vol, 1000.0, ang, 4, 0)
    np.testing.assert_allclose(a, b, rtol=2e-3, atol=2e-4)


def test_g3d_bp_axis_ray_weights():
    # one ray along +z through voxel-centre (x,y): floor/ceil give iL = i,
    # iU = i + 1, so interpXX = interpYY = 1 and each in-range sample adds
    # p * 0.5 * (interpZZ + interpZ) = p * 0.5 to the column
    M, N, p = 8, 10, 3.0
    i, j = 3, 5
    x = i + 0.5 - M / 2.0
    y = j + 0.5 - M / 2.0
    pts = np.zeros((1, 1, 6), np.float32, order="F")
    pts[0, 0] = [x, y, -50.0, x, y, 50.0]
    out = O.g3d_bp(np.full((1, 1, 1), p, np.float32, order="F"), np.zeros(1, np.float32), pts, M, N)
    assert np.count_nonzero(out.sum(axis=2)) == 1
    col = out[i, j].astype(np.float64)
    # samples at z in [0, N) every 0.5, skipped when ceil(z) >= N: 2N-1 of them
    assert col.sum() == pytest.approx(p * 0.5 * (2 * N - 1), rel=1e-6)


def test_g3d_bp_quarter_voxel_z_profile():
    # One ray along -x at y on a voxel centre (interpYY = 1) and z a quarter
    # voxel above a centre: kL = floor(z - bz) = 3, kU = 4, interpZ =
    # z - bz - kL + 0.5 = 0.75, interpZZ = 0.25.  It enters at x = M/2
    # exactly (SAD-free endpoints 12 and -4, alpha 1/2) and samples
    # X = x - bx = 8, 7.5, ..., 0.5.  A sample adds p*0.5*interpXX*interpZZ
    # to (iL, jL, kL) and p*0.5*interpXX*interpZ to (iL, jL, kU)
    # (CUDAmex_general3D_BP.cu:147-170), and is skipped when ceil(X) >= M
    # (X = 8 and 7.5).  interpXX = X - floor(X) + 0.5 per sample gives the
    # cell totals S = [1.0, 1.5 x 6, 0.5] along x.
    M, N, p = 8, 8, 3.0
    jL, kL = 5, 3
    y = jL + 0.5 - M / 2.0
    z = kL + 0.25 - N / 2.0
    pts = np.zeros((1, 1, 6), np.float32, order="F")
    pts[0, 0] = [12.0, y, z, -4.0, y, z]
    out = O.g3d_bp(np.full((1, 1, 1), p, np.float32, order="F"), np.zeros(1, np.float32), pts, M, N)
    S = np.array([1.0, 1.5, 1.5, 1.5, 1.5, 1.5, 1.5, 0.5])
    want = np.zeros((M, M, N))
    want[:, jL, kL] = p * 0.5 * 0.25 * S
    want[:, jL, kL + 1] = p * 0.5 * 0.75 * S
    np.testing.assert_array_equal(out, want.astype(np.float32))


# ------------------------------------------------------------------ OSC
def _osc_setup(W=16, H=8, n=8, seed=17):
    r = rng(seed)
    mu = f32(r.uniform(0.0, 0.02, size=(W, W, H)))
    pre = f32(np.full((W, H, n), 1e5))
    return mu, pre, (np.arange(n) * 2 * np.pi / n).astype(np.float32)


def test_osc_zero_volume_fixed_point_and_nonnegative():
    W, H, n = 16, 8, 8
    mu, pre, ang = _osc_setup(W, H, n)
    post = f32(pre * 0.5)
    for geom in (0, 1, 2):
        z, _, _ = O.osc_iter(np.zeros_like(mu), pre, post, 2, np.array([1, 0], np.int32), ang,
                             100.0, geom)
        assert not z.any()
        m2, _, _ = O.osc_iter(mu, pre, post, 2, np.array([1, 0], np.int32), ang, 100.0, geom)
        assert (m2 >= 0).all()


def test_osc_consistent_data_is_fixed_point():
    # post = pre*exp(-l) with l the OSC projection of mu makes num = 0
    W, H, n = 16, 8, 8
    mu, pre, ang = _osc_setup(W, H, n)
    for geom in (0, 1, 2):
        _, pr1, _ = O.osc_iter(mu, pre, np.zeros_like(pre), 1, np.array([0], np.int32), ang, 100.0,
                               geom)
        post = f32(pr1)  # = pre*exp(-l) - 0, bit for bit
        m2, pr2, _ = O.osc_iter(mu, pre, post, 1, np.array([0], np.int32), ang, 100.0, geom)
        assert not pr2.any()
        inside = _circle(W)
        np.testing.assert_array_equal(m2[inside], mu[inside])


def test_osc_outside_circle_untouched_and_views_once():
    W, H, n = 16, 6, 9
    mu, pre, ang = _osc_setup(W, H, n)
    post = f32(pre * 0.3)
    m2, pr2, po2 = O.osc_iter(mu, pre, post, 3, np.array([2, 0, 1], np.int32), ang, 100.0, 2)
    out = ~_circle(W)
    np.testing.assert_array_equal(m2[out], mu[out])
    assert (pr2 != pre).all() and (po2 != post).all()  # every view rewritten once


_libm = None


def _expf(x):
    """glibc expf, element by element (the oracle's exp of a float)."""
    global _libm
    if _libm is None:
        import ctypes

        _libm = ctypes.CDLL("libm.so.6")
        _libm.expf.restype = ctypes.c_float
        _libm.expf.argtypes = [ctypes.c_float]
    x = np.asarray(x, np.float32)
    return np.array([_libm.expf(float(v)) for v in x.ravel()], np.float32).reshape(x.shape)


def _std_geometry_pts(geom, W, H, sad):
    """The standard geometries' rays as general-3D pts [W,H,6] (beta = 0):
    source SAD along +x (on the axis for fan and cone, beside each detector
    pixel for parallel), flat detector SAD behind the axis, pixel (ix, iy)
    at u = ix - u0, v = iy - v0 times the magnification 2 across the fan
    (fan: u only; cone: u and v), u0 = (W-1)/2, v0 = (H-1)/2.  float32
    arithmetic in the order of the geometry definition
    (CUDAmex_oscIter.cu:150-190, CUDAmex_FP.cu:64-77,162-175,259-272)."""
    f = np.float32
    SR = RD = f(sad)
    u0, v0 = f(W - 1) / f(2), f(H - 1) / f(2)
    ix = np.arange(W, dtype=f)[:, None] * np.ones((1, H), f)
    iy = np.ones((W, 1), f) * np.arange(H, dtype=f)[None, :]
    zero = np.zeros((W, H), f)
    if geom == 0:
        u, v = ix - u0, iy - v0
        cols = (SR + zero, u, v, -RD + zero, u, v)
    elif geom == 1:
        u, v = (ix - u0) * (SR + RD) / SR, iy - v0
        cols = (SR + zero, zero, v, -RD + zero, u, v)
    else:
        u, v = (ix - u0) * (SR + RD) / SR, (iy - v0) * (SR + RD) / SR
        cols = (SR + zero, zero, zero, -RD + zero, u, v)
    return np.asfortranarray(np.stack(cols, axis=2).astype(f))


def _osc_project(vol, pre, post, ang, sad, geom):
    """One CUDAmex_oscIter subset over every view; returns (pre', post')."""
    _, pr, po = O.osc_iter(vol, pre, post, 1, np.array([0], np.int32), ang, sad, geom)
    return pr, po


@pytest.mark.parametrize("geom", [0, 1, 2])
def test_osc_projection_is_border_step_half_march(geom):
    # The OSC projection marches every geometry with step 0.5 and border
    # (zero) addressing (CUDAmex_oscIter.cu:244,419-425,460-462).  The
    # general-3D projector does exactly that for arbitrary rays, so with the
    # standard rays given as pts (same float operations after the rotation)
    # its ray sums are the OSC line integrals bit for bit.  The volume has
    # non-zero faces, so clamp addressing would differ; cone at step 1.0
    # would differ everywhere.
    W, H, n, sad = 16, 10, 5, 40.0
    r = rng(30 + geom)
    vol = f32(r.uniform(0.0, 0.05, size=(W, W, H)))
    ang = r.uniform(0, 2 * np.pi, n).astype(np.float32)
    pre = f32(r.uniform(1.0, 3.0, size=(W, H, n)))
    post = f32(r.uniform(0.0, 0.5, size=(W, H, n)))
    l = O.g3d_fp(vol, ang, _std_geometry_pts(geom, W, H, sad), n)
    e = _expf(-l)
    pr, po = _osc_project(vol, pre, post, ang, sad, geom)
    # the ratio terms of CUDAmex_oscIter.cu:266-268
    np.testing.assert_array_equal(pr, f32(pre * e - post))
    np.testing.assert_array_equal(po, f32(pre * e * l))


@pytest.mark.parametrize("geom", [0, 1])
def test_osc_projection_par_fan_equal_fp_on_zero_faces(geom):
    # par/fan CUDAmex_FP also steps 0.5; on a volume whose faces are zero
    # clamp and border fetch the same taps, so the OSC line integral is the
    # CUDAmex_FP ray sum bit for bit
    W, H, n, sad = 16, 10, 5, 40.0
    r = rng(33 + geom)
    vol = np.zeros((W, W, H), np.float32, order="F")
    vol[1:-1, 1:-1, 1:-1] = r.uniform(0.0, 0.05, size=(W - 2, W - 2, H - 2))
    ang = r.uniform(0, 2 * np.pi, n).astype(np.float32)
    pre = f32(np.ones((W, H, n)))
    pr, po = _osc_project(vol, pre, np.zeros_like(pre), ang, sad, geom)
    l = O.fp(vol, sad, ang, n, geom)
    np.testing.assert_array_equal(pr, _expf(-l))
    np.testing.assert_array_equal(po, f32(_expf(-l) * l))


def test_osc_cone_axis_ray_closed_form():
    # Cone, beta = 0, W and H odd, SAD a power of two: the central pixel's
    # ray runs along -x through texel centres in y and z and enters at
    # x = W/2 exactly.  With mu = 1/64 everywhere it takes 2W samples at
    # tex x = W, W-1/2, ..., 1/2; the first lerps T[W-1] with the border 0,
    # every other one sees 1/64: l = (W - 1/4)/64 exactly.  Clamp would
    # give W/64, step 1.0 (CUDAmex_FP's cone step) (W - 1/2)/64.
    W, H, sad = 15, 9, 512.0
    vol = f32(np.full((W, W, H), 1.0 / 64))
    ang = np.zeros(1, np.float32)
    pre = f32(np.ones((W, H, 1)))
    pr, po = _osc_project(vol, pre, np.zeros_like(pre), ang, sad, 2)
    l = np.float32((W - 0.25) / 64)
    assert pr[W // 2, H // 2, 0] == _expf(-l)
    assert po[W // 2, H // 2, 0] == np.float32(_expf(-l) * l)


def _embed1(P):
    """P [W,H,n] inside a zero frame one texel wide: [W+2,H+2,n]."""
    W, H, n = P.shape
    Q = np.zeros((W + 2, H + 2, n), np.float32, order="F")
    Q[1:-1, 1:-1] = P
    return Q


@pytest.mark.parametrize("geom", [0, 1, 2])
def test_osc_update_equals_bp_of_ratio_sinograms(geom):
    # num/den are the backprojections of pre'/post' with border addressing
    # and the weights of CUDAmex_oscIter.cu:320-355 (par U = UU = 1; fan
    # U = 1 for the row step, UU = SR/(SR-s) for the column and weight;
    # cone U = UU), which are CUDAmex_BP's geometries.  Border addressing of
    # P equals clamp addressing of P inside a one-texel zero frame, and the
    # (W+2, H+2) problem puts the same voxel at the same detector
    # coordinate + 1 (hw and v0 both grow by 1), so CUDAmex_BP of the
    # framed sinograms, cropped, is num/den.  The update is then
    # x <- max(x + x*num/den, 0) inside the circle, x unchanged outside
    # (CUDAmex_oscIter.cu:113-136).  The ratio sinograms have non-zero
    # edges and the fan/cone taps reach past them, so clamp in place of
    # border fails here, and so does a fan row step of UU.
    W, H, n, sad = 20, 10, 6, 30.0
    r = rng(40 + geom)
    mu = f32(r.uniform(0.001, 0.02, size=(W, W, H)))
    pre = f32(r.uniform(2.0, 3.0, size=(W, H, n)))
    post = f32(r.uniform(0.0, 1.0, size=(W, H, n)))
    ang = r.uniform(0, 2 * np.pi, n).astype(np.float32)
    mu2, pr, po = O.osc_iter(mu, pre, post, 1, np.array([0], np.int32), ang, sad, geom)
    num = O.bp(_embed1(pr), np.float32(sad), ang, geom)[1:-1, 1:-1, 1:-1].astype(np.float64)
    den = O.bp(_embed1(po), np.float32(sad), ang, geom)[1:-1, 1:-1, 1:-1].astype(np.float64)
    inside = _circle(W)
    m = mu.astype(np.float64)[inside]
    want = np.maximum(m + m * num[inside] / den[inside], 0.0)
    np.testing.assert_allclose(mu2[inside], want, rtol=2e-6, atol=0)
    np.testing.assert_array_equal(mu2[~inside], mu[~inside])


@pytest.mark.parametrize("geom", [0, 1, 2])
def test_osc_subsets_compose_in_pick_order(geom):
    # subset s takes views picks[s], picks[s] + S, picks[s] + 2S, ... and
    # the subsets run in pick order, each projecting the volume the previous
    # one left (CUDAmex_oscIter.cu:505-575): one call equals S calls of one
    # subset each on those views' sub-sinograms, bit for bit
    W, H, n, S, sad = 16, 8, 9, 3, 50.0
    r = rng(50 + geom)
    mu = f32(r.uniform(0.001, 0.02, size=(W, W, H)))
    pre = f32(r.uniform(2.0, 3.0, size=(W, H, n)))
    post = f32(r.uniform(0.0, 1.0, size=(W, H, n)))
    ang = r.uniform(0, 2 * np.pi, n).astype(np.float32)
    picks = np.array([2, 0, 1], np.int32)
    m_all, p_all, q_all = O.osc_iter(mu, pre, post, S, picks, ang, sad, geom)
    m, P, Q = mu, pre.copy(order="F"), post.copy(order="F")
    for pick in picks:
        views = np.arange(pick, n, S)
        m, p1, q1 = O.osc_iter(m, pre[:, :, views], post[:, :, views], 1, np.array([0], np.int32),
                               ang[views], sad, geom)
        P[:, :, views] = p1
        Q[:, :, views] = q1
    np.testing.assert_array_equal(m_all, m)
    np.testing.assert_array_equal(p_all, P)
    np.testing.assert_array_equal(q_all, Q)


def test_osc_f32_close_to_f64_twin():
    W, H, n = 16, 8, 8
    mu, pre, ang = _osc_setup(W, H, n)
    post = f32(pre * rng(18).uniform(0.3, 0.9, size=pre.shape))
    a, _, _ = O.osc_iter(mu, pre, post, 2, np.array([0, 1], np.int32), ang, 100.0, 2)
    b, _, _ = O.osc_iter(mu, pre, post, 2, np.array([0, 1], np.int32), ang, 100.0, 2, f64=True)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4


def _osc_tv_by_parts(pre_in, post_in, ang, sad, geom, K, tv, s0, s1, p, tv_steps, first_picks):
    """OSC_TV_recon.m:33-73 composed from the separately pinned oracle parts
    (subset schedule, CUDAmex_oscIter, CUDAmex_TVmin_3D) with the driver's
    own steps written out in numpy: mu starts at single(1e-10); pre/post are
    raised to 1e-6; from the second iteration on each call gets the raised
    data with element 1 zeroed (:41,62); TV runs only for k > 1 with
    c = TV * ||mu_old - mu||, mu_old's element 1 zeroed and NaN differences
    dropped (:45,64-68); NaN and Inf become 0 (:69-70)."""
    W, H, n = pre_in.shape
    mu = np.full((W, W, H), np.float32(1e-10), np.float32, order="F")
    eps = np.float32(1e-6)
    pre = f32(np.where(pre_in.astype(np.float64) < 1e-6, eps, pre_in))
    post = f32(np.where(post_in.astype(np.float64) < 1e-6, eps, post_in))
    pre2, post2 = pre.copy(order="F"), post.copy(order="F")
    pre2[0, 0, 0] = 0.0
    post2[0, 0, 0] = 0.0
    for k in range(1, K + 1):
        mu_old = mu.copy(order="F")
        mu_old[0, 0, 0] = 0.0
        ns = O.calc_num_subsets(s0, s1, p, K, k)
        picks = O.calc_subset_order(ns, int(first_picks[k - 1]))
        mu, _, _ = O.osc_iter(mu, pre, post, ns, picks, ang, sad, geom)
        pre, post = pre2.copy(order="F"), post2.copy(order="F")
        if k > 1:
            d = (mu_old - mu).astype(np.float64)
            d[np.isnan(d)] = 0.0
            d_a = np.float32(np.sqrt(np.cumsum((d * d).ravel(order="F"))[-1]))
            mu = O.tvmin(mu, np.float32(tv) * d_a, tv_steps)
            mu[~np.isfinite(mu)] = 0.0
    return mu


@pytest.mark.parametrize("geom", [1, 2])
def test_osc_tv_recon_is_composition_of_parts(geom):
    # The corner pixel of view 0 is far above the rest so that feeding it
    # back (not zeroing element 1 from the second iteration on) moves the
    # result; three iterations so that TV runs twice.
    W, H, n, sad = 16, 8, 12, 30.0
    r = rng(70 + geom)
    pre = f32(r.uniform(1e3, 2e3, size=(W, H, n)))
    post = f32(pre * r.uniform(0.4, 0.9, size=pre.shape))
    pre[0, 0, 0], post[0, 0, 0] = 5e4, 1e4
    post[3, 2, 1] = 0.0  # raised to 1e-6
    ang = (np.arange(n) * 2 * np.pi / n).astype(np.float32)
    fp = np.array([2, 1, 3], np.int32)
    args = (ang, sad, geom, 3, 0.1, 4, 2, 1.0, 5, fp)
    np.testing.assert_array_equal(O.osc_tv_recon(pre, post, *args), _osc_tv_by_parts(pre, post, *args))


def test_osc_tv_recon_f32_close_to_f64_twin():
    # the FP64 driver is the same template as the float one: a term dropped
    # or mistyped in either instantiation separates them far beyond rounding
    W, H, n = 16, 8, 12
    mu, pre, ang = _osc_setup(W, H, n)
    post = f32(pre * rng(21).uniform(0.3, 0.9, size=pre.shape))
    picks = np.array([1, 2, 1], np.int32)
    a = O.osc_tv_recon(pre, post, ang, 100.0, 2, 3, 0.1, 4, 2, 1.0, 5, picks)
    b = O.osc_tv_recon(pre, post, ang, 100.0, 2, 3, 0.1, 4, 2, 1.0, 5, picks, f64=True)
    assert b.dtype == np.float64
    e = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert 0 < e < 1e-4


# ------------------------------------------------------------------ TV
def test_tv_gradient_hand_computed_peak():
    # a single unit peak along x: g = 2/(1+eps) at the peak, -1/(1+eps) beside it
    x = np.zeros((7, 5, 5), np.float32, order="F")
    x[3, :, :] = 1.0
    g = O.tv_gradient(x)
    eps = np.float32(1e-6)
    one = np.float32(1.0) / (eps + np.float32(1.0))
    assert g[3, 2, 2] == pytest.approx(float(2 * one), rel=1e-7)
    assert g[2, 2, 2] == pytest.approx(float(-one), rel=1e-7)
    assert g[4, 2, 2] == pytest.approx(float(-one), rel=1e-7)
    assert g[1, 2, 2] == 0.0 and g[5, 2, 2] == 0.0
    assert not g[0].any() and not g[:, 0].any() and not g[:, :, 0].any()  # border


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_tv_gradient_step_edge_each_axis(axis):
    # x = 1 on and beyond plane p of `axis`, 0 before it: the first voxel of
    # the plane has only its backward difference along `axis` (g = 1/(1+eps)),
    # the voxel before it only its forward term along `axis` (g = -1/(1+eps))
    shape = [7, 6, 5]
    x = np.zeros(shape, np.float32, order="F")
    idx = [slice(None)] * 3
    idx[axis] = slice(3, None)
    x[tuple(idx)] = 1.0
    g = O.tv_gradient(x)
    one = float(np.float32(1.0) / (np.float32(1e-6) + np.float32(1.0)))
    at = lambda k: tuple(k if a == axis else 2 for a in range(3))  # noqa: E731
    assert g[at(3)] == pytest.approx(one, rel=1e-7)
    assert g[at(2)] == pytest.approx(-one, rel=1e-7)
    assert g[at(1)] == 0.0 and g[at(4)] == 0.0


def test_tv_gradient_is_gradient_of_tv_functional():
    # The gradient of CUDAmex_TVmin_3D.cu:101-158 is d/dx of the isotropic
    # TV with backward differences, TV(x) = sum_{v >= 1 on every axis}
    # sqrt(|x_v - x_{v-e1}|^2 + |x_v - x_{v-e2}|^2 + |x_v - x_{v-e3}|^2),
    # up to eps = 1e-6 in each denominator.  Differentiating that functional
    # (torch autograd, float64) checks every one of the 13 stencil taps,
    # including the six diagonal ones (xmypz, xpymz, ...) that only enter
    # the neighbours' norms.
    torch = pytest.importorskip("torch")
    r = rng(60)
    n = (9, 8, 7)
    base = np.cumsum(r.uniform(0.2, 1.0, size=n), axis=0)  # no flat spots
    x = f32(base + r.uniform(0.0, 0.6, size=n))
    t = torch.tensor(x.astype(np.float64), requires_grad=True)
    d0 = t[1:, 1:, 1:] - t[:-1, 1:, 1:]
    d1 = t[1:, 1:, 1:] - t[1:, :-1, 1:]
    d2 = t[1:, 1:, 1:] - t[1:, 1:, :-1]
    torch.sqrt(d0 * d0 + d1 * d1 + d2 * d2).sum().backward()
    want = t.grad.numpy()[1:-1, 1:-1, 1:-1]
    got = O.tv_gradient(x).astype(np.float64)[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.abs(want).max())


def test_tv_constant_volume_gives_nan():
    x = np.full((6, 6, 6), 0.5, np.float32, order="F")
    assert np.isnan(O.tvmin(x, 0.1, 1)).all()


def test_tv_step_length_is_c():
    x = f32(rng(19).uniform(size=(10, 9, 8)))
    c = 0.05
    y = O.tvmin(x, c, 1)
    assert np.linalg.norm(y.astype(np.float64) - x) == pytest.approx(c, rel=1e-4)
    y20 = O.tvmin(x, c, 20)
    assert np.linalg.norm(y20.astype(np.float64) - x) <= 20 * c * (1 + 1e-4)


# ------------------------------------------------------------------ FDK filter
def _kak_slaney_numpy(order):
    n = np.arange(order // 2 + 1)
    h = np.zeros(order // 2 + 1)
    h[0] = 0.25
    h[1::2] = -1.0 / (np.pi * n[1::2]) ** 2
    return np.concatenate([h, h[-2:0:-1]])


@pytest.mark.parametrize("window", ["ramp", "hamming", "shepp-logan", "cosine", "hann"])
@pytest.mark.parametrize("length", [20, 96, 384])
def test_design_filter_vs_numpy_fft(window, length):
    order = max(64, 1 << int(np.ceil(np.log2(2 * length))))
    f = 2 * np.real(np.fft.fft(_kak_slaney_numpy(order)))[: order // 2 + 1]
    w = 2 * np.pi * np.arange(f.size) / order
    d = 1.0
    win = {"ramp": np.ones_like(w),
           "hamming": .54 + .46 * np.cos(w / d),
           "cosine": np.cos(w / (2 * d)),
           "hann": (1 + np.cos(w / d)) / 2}
    if window == "shepp-logan":
        sl = np.ones_like(w)
        sl[1:] = np.sin(w[1:] / (2 * d)) / (w[1:] / (2 * d))
        win[window] = sl
    g = f.copy()
    g[1:] *= win[window][1:]
    full = np.concatenate([g, g[-2:0:-1]])
    got = O.design_filter(window, length)
    assert got.size == order
    np.testing.assert_allclose(got, full, rtol=1e-10, atol=1e-13)


def test_design_filter_dc_closed_form():
    order = 1024
    odd = np.arange(1, order // 2, 2, dtype=np.float64)
    dc = 2 * (0.25 - 2 * np.sum(1.0 / (np.pi * odd) ** 2))
    assert O.design_filter("ramp", 384)[0] == pytest.approx(dc, rel=1e-12)
    # the ramp's DC tends to 0 as order grows (sum 1/n^2 over odd n = pi^2/8)
    assert abs(dc) < 1e-3


@pytest.mark.parametrize("geom", ["par3d", "fan3d", "cone3d"])
def test_prefilter_f64_vs_numpy(geom):
    W, H, n, SAD = 40, 5, 4, 300.0
    sino = f32(rng(20).uniform(size=(W, H, n)))
    ang = (np.arange(n) * 0.25 + 0.1).astype(np.float32)
    got = O.prefilter(sino, geom, SAD, ang, "hamming", 1.0, f64=True)
    filt = O.design_filter("hamming", W)
    U = np.arange(1, W + 1)[:, None] - (W - 1) / 2.0
    V = np.arange(1, H + 1)[None, :] - (H - 1) / 2.0
    if geom == "cone3d":
        wgt = (SAD / np.sqrt(SAD ** 2 + U ** 2 + V ** 2)).astype(np.float32)
    elif geom == "fan3d":
        wgt = np.broadcast_to(SAD / np.sqrt(SAD ** 2 + U ** 2), (W, H)).astype(np.float32)
    else:
        wgt = np.ones((W, H), np.float32)
    for i in range(n):
        x = np.zeros((filt.size, H))
        x[:W] = (sino[:, :, i] * wgt).astype(np.float64)
        y = np.real(np.fft.ifft(np.fft.fft(x, axis=0) * filt[:, None], axis=0))[:W]
        dA = np.float32(ang[i] - ang[i - 1] if i > 0 else ang[1] - ang[0])
        np.testing.assert_allclose(got[:, :, i], y * float(dA) / 2, rtol=1e-9, atol=1e-12)


# ------------------------------------------------------------------ Parker
def _short_scan(W, sad, extra, n):
    delta = np.arctan((W / 2 - 0.5) / sad)
    return np.linspace(0, np.pi + 2 * delta + extra, n).astype(np.float32)


def test_parker_q1_is_parker_1982():
    # q = 1 on a scan of exactly pi + 2 delta (epsilon = 0) reduces to Parker's
    # closed form: sin^2(pi/4 b/(d-a)) rising, 1, sin^2(pi/4 (pi+2d-b)/(d+a))
    # falling, with a = -atan(u/SAD) for an increasing scan.
    W, sad = 64, 50.0
    ang = _short_scan(W, sad, 0.0, 400)
    win = O.parker_window(W, ang, "fan3d", sad, 1.0)
    al = -np.arctan((np.arange(W) - W / 2 + 0.5) / sad)[:, None]
    d = np.arctan((W / 2 - 0.5) / sad)
    B = np.abs(ang.astype(np.float64) - ang[0])[None, :]
    with np.errstate(all="ignore"):
        ref = np.ones((W, ang.size))
        r1 = B <= 2 * d - 2 * al
        r3 = B >= np.pi - 2 * al
        ref = np.where(r1, np.sin(np.pi / 4 * B / (d - al)) ** 2, ref)
        ref = np.where(r3, np.sin(np.pi / 4 * (np.pi + 2 * d - B) / (d + al)) ** 2, ref)
    ok = np.isfinite(ref)
    assert (~ok).sum() <= 1  # the edge ray at beta = 0, 0/0 in both
    np.testing.assert_allclose(win[ok], ref[ok], atol=1e-5)
    assert win.min() >= 0 and win.max() <= 1


@pytest.mark.parametrize("q", [1.0, 0.5, 0.1])
@pytest.mark.parametrize("geom", ["fan3d", "par3d"])
def test_parker_conjugate_rays_sum_to_one(q, geom):
    # Wesarg's weights are normalised: the ray (beta, a) and its conjugate
    # (beta + pi + 2a, -a) -- column W-1-u -- carry weights summing to 1,
    # also with an over-scan (epsilon > 0).
    W, sad = 64, 50.0
    ext = float(_short_scan(W, sad, 0.3, 2)[-1])
    al = -np.arctan((np.arange(W) - W / 2 + 0.5) / sad) if geom == "fan3d" else np.zeros(W)
    r = rng(31)
    done = 0
    while done < 60:
        u = int(r.integers(1, W - 1))
        b0 = float(r.uniform(0, ext))
        b1 = b0 + np.pi + 2 * al[u]
        if not 0 < b1 < ext:
            continue
        ang = np.array([0.0, b0, b1, ext], np.float32)
        w = O.parker_window(W, ang, geom, sad, q)
        assert abs(w[u, 1] + w[W - 1 - u, 2] - 1) < 1e-5
        done += 1


@pytest.mark.parametrize("geom", ["fan3d", "cone3d"])
def test_parker_short_scan_fdk(geom):
    # a Parker-weighted short scan reconstructs as well as the full scan (whose
    # filter counts every ray twice, so it is compared at half scale), clearly
    # better than without the weights; the reversed scan direction (alpha
    # not flipped, negative angle steps) gives exactly the negated volume.
    W = H = 48
    sad = 80.0
    g = O.GEOM_TYPES[geom]
    c = (W - 1) / 2.0
    x, y, z = np.meshgrid(np.arange(W) - c, np.arange(W) - c, np.arange(H) - c, indexing="ij")
    vol = f32(0.02 * ((x / 16) ** 2 + (y / 12) ** 2 + (z / 14) ** 2 <= 1)
              + 0.01 * (((x - 5) / 4) ** 2 + ((y + 3) / 5) ** 2 + (z / 6) ** 2 <= 1))
    full = np.linspace(0, 2 * np.pi, 241)[:-1].astype(np.float32)
    short = _short_scan(W, sad, 0.05, 130)

    def rec(ang, q):
        P = O.fp(vol, sad, ang, ang.size, g)
        return O.bp(O.prefilter(P, geom, sad, ang, "hamming", 1.0, parker_q=q), sad, ang, g)

    mid = slice(H // 2 - 3, H // 2 + 3)

    def err(v):
        return np.linalg.norm(v[:, :, mid] - vol[:, :, mid]) / np.linalg.norm(vol[:, :, mid])
    e_full = err(rec(full, None) / 2)
    rp = rec(short, 1.0)
    assert err(rp) < 1.15 * e_full
    assert err(rec(short, None)) > 2 * err(rp)
    rr = rec(short[::-1].copy(), 1.0)
    assert np.abs(rr + rp).max() <= 1e-5 * np.abs(rp).max()


def test_prefilter_parker_is_weighted_sinogram():
    # parker_q applies the window after the cosine weight, before filtering
    W, H, n, sad = 32, 6, 40, 60.0
    sino = f32(rng(41).uniform(size=(W, H, n)))
    ang = _short_scan(W, sad, 0.1, n)
    win = O.parker_window(W, ang, "cone3d", sad, 0.1)
    U = np.arange(1, W + 1)[:, None] - (W - 1) / 2.0
    V = np.arange(1, H + 1)[None, :] - (H - 1) / 2.0
    wgt = (sad / np.sqrt(sad ** 2 + U ** 2 + V ** 2)).astype(np.float32)
    pre = f32((sino * wgt[:, :, None]) * win[:, None, :])
    a = O.prefilter(sino, "cone3d", sad, ang, "hamming", 1.0, parker_q=0.1)
    # the weighted sinogram through the par3d branch (no cosine weight, no Parker)
    b = O.prefilter(pre, "par3d", sad, ang, "hamming", 1.0)
    assert np.array_equal(a, b)


def test_log_transform_vs_numpy():
    r = rng(43)
    pre = r.uniform(0.5, 2.0, 1000).astype(np.float32)
    post = r.uniform(0.0, 2.0, 1000).astype(np.float32)
    pre[:5] = [0, 0, 1, -1, np.inf]
    post[:5] = [0, 1, 0, 1, 1]
    with np.errstate(all="ignore"):
        ref = -np.log(post / pre)
    ref[~np.isfinite(ref)] = 0
    got = O.log_transform(pre, post)
    np.testing.assert_allclose(got, ref, rtol=2e-7, atol=0)
    assert not got[:5].any()


# ------------------------------------------------------------------ subsets
def test_calc_num_subsets_closed_form():
    s0, s1, p, K = 64, 2, 0.2, 10
    for k in range(1, K + 1):
        v = (s0 - s1) / (K - 1) ** p * (K - k) ** p + s1
        expect = int(np.floor(v + 0.5))
        assert O.calc_num_subsets(s0, s1, p, K, k) == expect
    assert O.calc_num_subsets(64, 2, 0.2, 10, 1) == 64
    assert O.calc_num_subsets(64, 2, 0.2, 10, 10) == 2


def test_calc_subset_order_properties():
    for n in range(1, 30):
        for first in range(1, n + 1):
            picks = O.calc_subset_order(n, first)
            assert sorted(picks.tolist()) == list(range(n))
            assert picks[0] == first - 1
            remaining = [v for v in range(1, n + 1) if v != first]
            cur = first
            for nxt in picks[1:] + 1:
                lo = [v for v in remaining if v < cur]
                hi = [v for v in remaining if v > cur]
                side = lo if len(lo) > len(hi) else hi
                assert nxt == side[(len(side) + 1) // 2 - 1]
                remaining.remove(nxt)
                cur = nxt


def test_subset_order_known_case():
    # n = 10 from pick 3: the middle of the larger side each time
    assert O.calc_subset_order(10, 3).tolist() == [2, 6, 3, 7, 1, 5, 8, 0, 4, 9]
