"""Parity of the HIP operators with the CPU oracle on seeded inputs.

Plain FP32 must match the oracle within 1e-4 relative max and 1e-6
relative L2 (FP, BP); the standard-geometry FP and BP are expected to be
bit-identical and a separate test checks that.  Every call goes through the
C ABI (libmexct.so) via the mexct package.
"""
import numpy as np
import pytest

import mexct
import oracle as O
from conftest import REL_L2, REL_MAX, rel_errors

pytestmark = pytest.mark.gpu

GEOMS = [0, 1, 2]


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _vol(W, H, seed, lo=0.0, hi=1.0):
    return np.asfortranarray(_rng(seed).uniform(lo, hi, size=(W, W, H)).astype(np.float32))


def _angles(n, span=2 * np.pi, off=0.1):
    return ((np.arange(n) * span / n) + off).astype(np.float32)


def _check(a, b, what, rel_max=REL_MAX, rel_l2=REL_L2):
    m, l2 = rel_errors(a, b)
    print(f"{what}: rel_max={m:.3e} rel_l2={l2:.3e} exact={np.mean(a == b):.6f}")
    assert m <= rel_max and l2 <= rel_l2, f"{what}: rel_max={m:.3e} rel_l2={l2:.3e}"


@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("sad", [60.0, 1000.0])
def test_fp_std(gpu, geom, sad):
    W, H, n = 40, 36, 10
    vol = _vol(W, H, 11 + geom)
    ang = _angles(n)
    got = mexct.CUDAmex_FP(vol, sad, ang, n, geom)
    ref = O.fp(vol, sad, ang, n, geom)
    _check(got, ref, f"FP geom={geom} sad={sad}")
    assert np.array_equal(got, ref), "FP is expected to be bit-identical to the oracle"


def test_fp_unwritten_rows(gpu):
    # rows iy >= 32*ceil(W/32) stay 0 (reference grid, CUDAmex_FP.cu:432-440)
    W, H, n = 20, 50, 3
    vol = _vol(W, H, 5)
    ang = _angles(n)
    for geom in GEOMS:
        got = mexct.CUDAmex_FP(vol, 500.0, ang, n, geom)
        ref = O.fp(vol, 500.0, ang, n, geom)
        assert np.array_equal(got, ref)
        assert not got[:, 32:, :].any() and got[:, :32, :].any()


@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("zrange", [32, 48, 64])
@pytest.mark.parametrize("sad", [60.0, 1000.0])
def test_bp_std(gpu, geom, zrange, sad):
    W, H, n = 40, 72, 9
    sino = np.asfortranarray(_rng(21 + geom).normal(size=(W, H, n)).astype(np.float32))
    ang = _angles(n)
    got = mexct.CUDAmex_BP(sino, sad, ang, geom, zrange=zrange)
    ref = O.bp(sino, sad, ang, geom)
    _check(got, ref, f"BP geom={geom} zr={zrange} sad={sad}")
    assert np.array_equal(got, ref), "BP is expected to be bit-identical to the oracle"


@pytest.mark.parametrize("ratio", [1.02, 1.3, 1.8, 2.5, 10.0])
@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
def test_bp_cone_source_near_circle(gpu, ratio, storage):
    # SAD = ratio x (hw + 1): at 1.02 and 1.3 the detector rows of the circle's
    # voxels reach far past the detector and the kernel clamps them; at 1.8
    # and 2.5 a halo covering them would add more than a third of the rows,
    # so the kernel still clamps (the f16 row-pair and row-quad paths); at 10
    # the f16 sinogram copy's row halo covers them and no clamp runs (f32
    # always clamps).  Either way the result is the reference's clamp
    # addressing.
    W, H, n = 40, 30, 9
    sad = ratio * (W // 2)
    sino = np.asfortranarray(_rng(77).uniform(0, 1, (W, H, n)).astype(np.float32))
    ang = _angles(n)
    got = mexct.CUDAmex_BP(sino, sad, ang, 2, storage=storage)
    if storage == mexct.F32:
        assert np.array_equal(got, O.bp(sino, sad, ang, 2)), f"sad={sad}"
    else:  # against the oracle on the f16-rounded sinogram
        ref = O.bp(_round_storage(sino, storage), sad, ang, 2)
        # below 2.02 hw the magnification reaches 1 / (ratio - 1) at the
        # source side (50 at 1.02), where a few heavily weighted samples make
        # up each voxel's sum and the f16 lerp rounding averages out less
        tol = F16_LERP_BOUND if ratio > 2.02 else (1e-3, 5e-4)
        _check(got, ref, f"BP f16 sad={sad} vs rounded", *tol)


def test_bp_odd_width(gpu):
    W, H, n = 33, 17, 5
    sino = np.asfortranarray(_rng(3).normal(size=(W, H, n)).astype(np.float32))
    ang = _angles(n)
    for geom in GEOMS:
        assert np.array_equal(mexct.CUDAmex_BP(sino, 300.0, ang, geom), O.bp(sino, 300.0, ang, geom))


def _random_pts(W, H, seed, spread=40.0):
    r = _rng(seed)
    a = r.uniform(-spread, spread, size=(W, H, 3))
    b = r.uniform(-spread, spread, size=(W, H, 3))
    return np.asfortranarray(np.concatenate([a, b], axis=2).astype(np.float32))


def test_g3d_fp_shared_pts(gpu):
    side, nsl, n = 24, 20, 5
    vol = _vol(side, nsl, 31)
    pts = _random_pts(16, 12, 32)
    ang = _angles(n)
    got = mexct.CUDAmex_general3D_FP(vol, ang, pts, n)
    ref = O.g3d_fp(vol, ang, pts, n)
    _check(got, ref, "g3D FP shared")
    assert np.array_equal(got, ref)


def test_g3d_fp_per_view_pts(gpu):
    side, nsl, n = 24, 20, 4
    vol = _vol(side, nsl, 33)
    pts = np.asfortranarray(np.stack([_random_pts(10, 14, 40 + k) for k in range(n)], axis=3))
    ang = np.zeros(n, np.float32)
    got = mexct.CUDAmex_general3D_FP(vol, ang, pts, n)
    ref = O.g3d_fp(vol, ang, pts, n)
    assert np.array_equal(got, ref)


def test_g3d_fp_matches_parallel(gpu):
    # parallel rays given as pts (general3D_projection_test.m:13-29) equal
    # CUDAmex_FP par at axis-aligned views, where both sample sets sit on
    # the same half-voxel grid; the volume edge is zero so clamp and border
    # addressing agree.
    from mexct import phantoms
    W, H = 32, 24
    vol = np.zeros((W, W, H), np.float32, order="F")
    vol[4:-4, 4:-4, 4:-4] = _rng(7).uniform(size=(W - 8, W - 8, H - 8))
    ang = (np.arange(4) * np.pi / 2).astype(np.float32)
    a = mexct.CUDAmex_general3D_FP(vol, ang, phantoms.parallel_pts(W, H), 4)
    b = mexct.CUDAmex_FP(vol, 1000.0, ang, 4, 0)
    m, l2 = rel_errors(a, b)
    print(f"g3D parallel vs FP par: rel_max={m:.3e} rel_l2={l2:.3e}")
    assert m < 1e-3 and l2 < 1e-4


def test_g3d_bp(gpu):
    W, H, n, M, N = 12, 10, 6, 20, 16
    sino = np.asfortranarray(_rng(51).uniform(size=(W, H, n)).astype(np.float32))
    pts = _random_pts(W, H, 52, spread=25.0)
    ang = _angles(n)
    got = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N)
    ref = O.g3d_bp(sino, ang, pts, M, N)
    # float atomics: the summation order differs from the oracle's canonical order
    _check(got, ref, "g3D BP", rel_max=1e-5, rel_l2=1e-6)


def _g3d_bp_case():
    # per-view random rays through and around a 37 x 37 x 45 volume (no
    # multiple of the 16 x 16 x 32 bricks): many cross it, some miss it, and
    # one per view ends 400 voxels short of it, so that its walk back to the
    # box face is far longer than the volume's diagonal (the atomic kernel
    # takes such rays)
    W, H, n, M, N = 20, 12, 5, 37, 45
    sino = np.asfortranarray(_rng(61).uniform(size=(W, H, n)).astype(np.float32))
    pts = np.stack([_random_pts(W, H, 62 + k, spread=35.0) for k in range(n)], axis=3)
    for k in range(n):
        pts[0, 0, :, k] = [-1000.0, 0.5 + k, 0.25, -420.0, 0.5 + k, 0.25]
    return sino, (np.arange(n) * 0.7).astype(np.float32), np.asfortranarray(pts), M, N


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_g3d_bp_modes(gpu, mode):
    sino, ang, pts, M, N = _g3d_bp_case()
    got = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=mode)
    ref = O.g3d_bp(sino, ang, pts, M, N)
    # float sums in another order than the oracle's canonical one (modes 0
    # and 2), or exact fixed-point sums rounded once (mode 1)
    _check(got, ref, f"g3D BP mode {mode}", rel_max=1e-5, rel_l2=1e-6)


def test_g3d_bp_deterministic_bits_any_chunking(gpu, monkeypatch):
    # fixed-point sums are exact, so the bits depend neither on the run nor
    # on how the views are cut into chunks; the float bricks split in chunks
    # stay within the tolerance
    sino, ang, pts, M, N = _g3d_bp_case()
    a = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=1)
    b = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=1)
    assert np.array_equal(a, b)
    monkeypatch.setenv("MEXCT_G3D_CHUNK_VIEWS", "2")
    c = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=1)
    assert np.array_equal(a, c)
    f = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=0)
    _check(f, O.g3d_bp(sino, ang, pts, M, N), "g3D BP bricks, 2-view chunks", 1e-5, 1e-6)


def _boundary_rays():
    """Rays on and near the 16 x 16 x 12 brick faces of a 48 x 48 x 36
    volume (world x, y = -8 and 8, z = -6 and 6 are brick faces; -24 / 24
    and -18 / 18 the volume's): along the faces and their edges, through a
    brick corner, with slopes so small that they cross a face only after
    a hundred samples or more, starting or ending inside the volume, and
    grazing the volume's own faces; each also reversed.  The brick pass
    marks bricks from the exact line widened by the running sum's drift and
    skips the cell test where the line is that far inside a brick: these
    are the rays such bounds must still cover."""
    e = 1e-4
    rays = [
        (-40, -8, -6, 40, -8, -6), (-40, -8 + e, -6, 40, -8 + e, -6), (-40, -8, 6 - e, 40, -8, 6 - e),
        (8, -40, 0, 8, 40, 0), (8 - e, -40, 6, 8 - e, 40, 6), (-8, 8, -30, -8, 8, 30),
        (-8 + e, 8 - e, -30, -8 + e, 8 - e, 30), (-40, -8.001, 0.5, 40, -7.999, 0.5),
        (-40, 3.0, 5.999, 40, 3.0, 6.001), (-40, -7.9995, -6.0005, 40, -8.0005, -5.9995),
        (-30, -30, -22.5, 30, 30, 22.5), (-30, 30, 22.5, 30, -30, -22.5), (0, 0, 0, 50, 3, 1),
        (1, -2, 3, -60, 5, -2), (-40, -24, 0, 40, -24, 0), (-40, 0, 18, 40, 0, 18),
        (-40, 23.9999, -17.9999, 40, 23.9999, -17.9999), (-24, -40, 5, -24, 40, 5),
        (-40, 16, -6, 40, 16.0002, -6), (-40, -8, 12, 40, 8, 12.00001),
    ]
    rays += [(r[3], r[4], r[5], r[0], r[1], r[2]) for r in rays]
    return np.array(rays, np.float64)


def test_g3d_bp_brick_boundary_rays(gpu):
    W, H, M, N = 8, 5, 48, 36  # 40 rays per view
    rays = _boundary_rays()
    rng = _rng(71)
    n = 2
    pts = np.zeros((W, H, 6, n), np.float32)
    for k in range(n):
        # both views hold every boundary ray, in another order, then random
        # rays near the faces
        order = rng.permutation(len(rays))
        for q in range(W * H):
            if q < len(rays):
                r = rays[order[q]]
            else:
                a = rng.choice([-8.0, 8.0, -6.0, 6.0]) + rng.uniform(-1e-3, 1e-3, 6)
                r = np.array([-40, a[0], a[1], 40, a[2], a[3]])
            pts[q % W, q // W, :, k] = r
    pts = np.asfortranarray(pts)
    sino = np.asfortranarray(rng.uniform(0.5, 1.5, size=(W, H, n)).astype(np.float32))
    ang = np.zeros(n, np.float32)  # no rotation: the rays lie exactly where written
    ref = O.g3d_bp(sino, ang, pts, M, N)
    assert np.count_nonzero(ref) > 0
    for mode in (0, 2):
        got = mexct.CUDAmex_general3D_BP(sino, ang, pts, M, N, mode=mode)
        _check(got, ref, f"g3D BP boundary rays mode {mode}", 1e-5, 1e-6)


def _osc_inputs(W, H, n, seed):
    r = _rng(seed)
    mu = np.asfortranarray(r.uniform(0.0, 0.03, size=(W, W, H)).astype(np.float32))
    pre = np.asfortranarray(r.uniform(5e4, 1e5, size=(W, H, n)).astype(np.float32))
    post = np.asfortranarray((pre * r.uniform(0.2, 0.9, size=(W, H, n))).astype(np.float32))
    return mu, pre, post


# the second shape: many z-ranges of the backprojection (the cone row start
# replays the running sum over whole ranges), a last range cut short, and
# tiles cut by the volume edge
@pytest.mark.parametrize("W,H,n,ns", [(32, 20, 12, 4), (26, 70, 8, 2)])
@pytest.mark.parametrize("geom", GEOMS)
def test_osc_iter(gpu, geom, W, H, n, ns):
    mu, pre, post = _osc_inputs(W, H, n, 61 + geom)
    picks = np.array([2, 0, 3, 1] if ns == 4 else [1, 0], np.int32)
    ang = _angles(n)
    rmu, rpre, rpost = O.osc_iter(mu, pre, post, ns, picks, ang, 200.0, geom)
    mexct.CUDAmex_oscIter(mu, pre, post, ns, picks, ang, 200.0, geom)
    # expf on the device and in glibc may differ by an ulp
    _check(pre, rpre, f"OSC pre geom={geom}", 1e-5, 1e-6)
    _check(post, rpost, f"OSC post geom={geom}", 1e-5, 1e-6)
    _check(mu, rmu, f"OSC mu geom={geom}", 1e-4, 1e-5)


def test_osc_empty_subset_zeroes_circle(gpu):
    # a pick >= nProj selects no view: num = den = 0 and the update gives 0
    W, H, n = 16, 8, 4
    mu, pre, post = _osc_inputs(W, H, n, 9)
    picks = np.array([7], np.int32)
    ang = _angles(n)
    rmu, _, _ = O.osc_iter(mu, pre, post, 1, picks, ang, 100.0, 2)
    mexct.CUDAmex_oscIter(mu, pre, post, 1, picks, ang, 100.0, 2)
    assert np.array_equal(mu, rmu)


# the gradient kernel reduces 64 x 4 x 4 voxels per workgroup: shapes with
# partial workgroups along every axis, and a depth below one workgroup's 4
@pytest.mark.parametrize("shape", [(20, 18, 16), (70, 9, 7), (33, 5, 3)])
def test_tvmin(gpu, shape):
    x = np.asfortranarray(_rng(71).uniform(size=shape).astype(np.float32))
    got = mexct.CUDAmex_TVmin_3D(x, 0.05)
    ref = O.tvmin(x, 0.05, 20)
    _check(got, ref, "TVmin", 1e-5, 1e-6)


# The steps of one call keep their state (gradient, norm, step constant) on
# the device: the result must be the bits of the same steps taken one call at
# a time, for odd and even counts.
@pytest.mark.parametrize("steps", [2, 3, 6])
def test_tvmin_steps_equal_single_steps(gpu, steps):
    x = np.asfortranarray(_rng(72).uniform(size=(70, 9, 7)).astype(np.float32))
    fused = mexct.CUDAmex_TVmin_3D(x, 0.05, steps=steps)
    one = x
    for _ in range(steps):
        one = mexct.CUDAmex_TVmin_3D(one, 0.05, steps=1)
    assert np.array_equal(fused, one)


@pytest.mark.parametrize("geom", ["par3d", "fan3d", "cone3d"])
@pytest.mark.parametrize("window", ["ramp", "hamming", "shepp-logan"])
def test_prefilter(gpu, geom, window):
    W, H, n = 48, 6, 5
    sino = np.asfortranarray(_rng(81).uniform(size=(W, H, n)).astype(np.float32))
    ang = _angles(n)
    got = mexct.preFilterSinogram(sino, {"type": geom, "SAD": 500.0}, ang, window, 1.0)
    ref = O.prefilter(sino, geom, 500.0, ang, window, 1.0)
    _check(got, ref, f"prefilter {geom} {window}", 1e-6, 1e-7)


def test_fp_long_ray_compensated_sum(gpu):
    # beta = 0 parallel rays run along x through texel centres and midpoints
    # (0.5 steps), so with zero x-faces each sample of an F16-exact volume is
    # exact and the ray sum is the column sum, known exactly in FP64 (every
    # other x texel is 0, so each packed-f16 x lerp halves one value, exact
    # in f16 too, while the values span nine binades).  The
    # 2048-sample rays exceed the default long-ray bound of 1024 samples: the
    # compensated sum stays within a few ulps, the plain running sum (long_ray
    # < 0) drifts with the sample count.
    W, H = 1024, 2
    r = _rng(85)
    vol = np.zeros((W, W, H), np.float32, order="F")
    vals = r.uniform(0.5, 1.0, (W - 2, W, H)) * np.exp2(-r.integers(0, 9, (W - 2, W, H)))
    vals[::2] = 0.0
    vol[1:-1] = vals.astype(np.float16).astype(np.float32)
    ang = np.zeros(1, np.float32)
    exact = vol.astype(np.float64).sum(axis=0)
    comp = mexct.CUDAmex_FP(vol, 1000.0, ang, 1, 0, storage=mexct.F16)[:, :, 0]
    plain = mexct.CUDAmex_FP(vol, 1000.0, ang, 1, 0, storage=mexct.F16, long_ray=-1)[:, :, 0]
    e_comp = np.abs(comp - exact).max() / np.abs(exact).max()
    e_plain = np.abs(plain - exact).max() / np.abs(exact).max()
    print(f"2048-sample rays: compensated rel err {e_comp:.3e}, plain {e_plain:.3e}")
    assert e_comp <= 3e-7
    assert e_plain > 3 * e_comp
    # a short-ray launch takes the plain loop whatever long_ray says
    small = np.asfortranarray(vol[:64, :64, :])
    a = mexct.CUDAmex_FP(small, 1000.0, ang, 1, 0, storage=mexct.F16)
    b = mexct.CUDAmex_FP(small, 1000.0, ang, 1, 0, storage=mexct.F16, long_ray=-1)
    assert np.array_equal(a, b)


def _short_scan(W, sad, extra, n, geom):
    delta = np.arctan((W / 2 - 0.5) / sad) if geom != "par3d" else 0.0
    return np.linspace(0, np.pi + 2 * delta + extra, n).astype(np.float32)


@pytest.mark.parametrize("geom", ["par3d", "fan3d", "cone3d"])
@pytest.mark.parametrize("q", [1.0, 0.1])
def test_prefilter_parker(gpu, geom, q):
    W, H, n, sad = 48, 6, 40, 80.0
    sino = np.asfortranarray(_rng(82).uniform(size=(W, H, n)).astype(np.float32))
    ang = _short_scan(W, sad, 0.1, n, geom)
    got = mexct.preFilterSinogram(sino, {"type": geom, "SAD": sad}, ang, "hamming", 1.0,
                                  parker_q=q)
    ref = O.prefilter(sino, geom, sad, ang, "hamming", 1.0, parker_q=q)
    _check(got, ref, f"prefilter+Parker {geom} q={q}", 1e-6, 1e-7)


@pytest.mark.parametrize("case", ["fan-up", "fan-down", "cone-exact", "par-over", "longer-angles"])
def test_parker_weight(gpu, case):
    # the window on a sinogram of ones: device sinf against glibc's, within
    # an ulp or two of the window's [0, 1] range.  fan-down runs the angles
    # backwards (alpha not flipped); cone-exact has epsilon = 0 and the 0/0
    # edge ray; longer-angles passes more angles than views (extent and
    # direction from all of them).
    W, H, sad = 64, 3, 50.0
    geom = {"fan-up": "fan3d", "fan-down": "fan3d", "cone-exact": "cone3d",
            "par-over": "par3d", "longer-angles": "fan3d"}[case]
    ang = _short_scan(W, sad, 0.0 if case == "cone-exact" else 0.3, 120, geom)
    if case == "fan-down":
        ang = ang[::-1].copy()
    n = 100 if case == "longer-angles" else ang.size
    sino = np.asfortranarray(np.ones((W, H, n), np.float32))
    got = mexct.parker_weight(sino, ang, {"type": geom, "SAD": sad}, 0.1)
    win = O.parker_window(W, ang, geom, sad, 0.1)[:, :n]
    ref = np.broadcast_to(win[:, None, :], (W, H, n))
    d = np.abs(got - ref).max()
    print(f"Parker window {case}: max|diff|={d:.3e} exact={np.mean(got == ref):.6f}")
    assert d <= 3e-7
    assert np.array_equal(got[:, 0, :], got[:, 2, :])
    assert 0.02 < win.mean() < 0.98  # a real window, not all 0 or 1


@pytest.mark.parametrize("geom", ["par3d", "fan3d", "cone3d"])
@pytest.mark.parametrize("q", [None, 0.1])
def test_fbp_recon(gpu, geom, q):
    # FBP_recon: log transform (with zero, negative and infinite ratios),
    # pre-filter with or without Parker, then the BP of the geometry
    W, H, sad = 40, 8, 120.0
    ang = _short_scan(W, sad, 0.2, 60, geom) if q is not None else _angles(60)
    r = _rng(83)
    pre = r.uniform(0.8, 1.2, (W, H, ang.size)).astype(np.float32)
    post = (pre * np.exp(-r.uniform(0, 2, pre.shape))).astype(np.float32)
    post[0, 0, :3] = 0          # -log(0) = Inf -> 0
    pre[1, 0, :3] = 0           # x/0 -> -Inf -> 0
    post[2, 1, :3] = -1         # negative ratio -> 0
    pre = np.asfortranarray(pre)
    post = np.asfortranarray(post)
    g = {"type": geom, "SAD": sad}
    got = mexct.FBP_recon(pre, post, ang, g, "hamming", 1.0, parker_q=q)
    ref = O.fbp_recon(pre, post, ang, geom, sad, "hamming", 1.0, parker_q=q)
    _check(got, ref, f"FBP_recon {geom} q={q}", 1e-5, 1e-6)


def test_fbp_recon_device_context(gpu):
    # the _dev entry on torch tensors, with F16 BP storage: against the host
    # entry's F32 result on the same filtered data
    import torch
    from mexct import device
    W, H, sad = 48, 16, 200.0
    ang = _short_scan(W, sad, 0.1, 90, "cone3d")
    r = _rng(84)
    pre = np.asfortranarray(r.uniform(0.9, 1.1, (W, H, ang.size)).astype(np.float32))
    post = np.asfortranarray((pre * np.exp(-r.uniform(0, 1, pre.shape))).astype(np.float32))
    ref = mexct.FBP_recon(pre, post, ang, {"type": "cone3d", "SAD": sad}, "hamming", 1.0,
                          parker_q=1.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda()
    ctx = device.Context(0)
    out = torch.zeros((H, W, W), dtype=torch.float32, device="cuda")
    ctx.fbp_recon(t(pre), t(post), ang, 2, sad, 3, 1.0, out, parker_q=1.0)
    torch.cuda.synchronize()
    got = out.cpu().numpy().transpose(2, 1, 0)
    _check(got, ref, "FBP_recon _dev f32", 0.0, 0.0)
    ctx.fbp_recon(t(pre), t(post), ang, 2, sad, 3, 1.0, out, parker_q=1.0, storage=mexct.F16)
    torch.cuda.synchronize()
    _check(out.cpu().numpy().transpose(2, 1, 0), ref, "FBP_recon _dev f16", 2e-3, 1e-3)
    ctx.close()


def test_osc_tv_recon(gpu):
    W, H, n = 24, 12, 12
    from mexct import phantoms
    vol = phantoms.osc_phantom(W, seed=3)[:, :, :H].copy(order="F")
    ang = _angles(n)
    fp = O.fp(vol, 300.0, ang, n, 2)
    pre, post = phantoms.poisson_transmission(fp, 1e5, seed=33)
    picks = mexct.first_picks(5, 4, 2, 0.5, 3)
    got = mexct.OSC_TV_recon(pre, post, ang, {"type": "cone3d", "SAD": 300.0}, 3, 0.1, 4, 2, 0.5,
                             tv_steps=10, picks=picks)
    ref = O.osc_tv_recon(pre, post, ang, 300.0, 2, 3, 0.1, 4, 2, 0.5, 10, picks)
    _check(got, ref, "OSC-TV", 1e-4, 1e-5)


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16])
def test_osc_tv_resume_equals_uninterrupted(gpu, storage):
    # checkpointing after iterations 1 and 3 of a 5-iteration schedule that
    # changes its subset count every iteration (6 down to 2) and resuming
    # from (mu, k) gives the bits of the uninterrupted run
    W, H, n = 32, 16, 24
    from mexct import phantoms
    vol = phantoms.osc_phantom(W, seed=3)[:, :, :H].copy(order="F")
    ang = _angles(n)
    pre, post = phantoms.poisson_transmission(O.fp(vol, 300.0, ang, n, 2), 1e5, seed=33)
    picks = mexct.first_picks(9, 6, 2, 1.0, 5)
    args = (pre, post, ang, {"type": "cone3d", "SAD": 300.0}, 5, 0.1, 6, 2, 1.0)
    kw = dict(tv_steps=4, picks=picks, storage=storage)
    whole = mexct.OSC_TV_recon(*args, **kw)
    mu = mexct.OSC_TV_recon(*args, iterations=(1, 1), **kw)
    mu = mexct.OSC_TV_recon(*args, iterations=(2, 3), mu=mu, **kw)
    mu = mexct.OSC_TV_recon(*args, iterations=(4, 5), mu=mu, **kw)
    assert np.array_equal(mu, whole)
    assert not np.array_equal(mexct.OSC_TV_recon(*args, iterations=(1, 3), **kw), whole)


# relative RMSE bound vs the FP64 oracle: FP16 storage (11-bit significand)
# must stay <= 1e-3; BF16 (8-bit significand) rounds a constant 0.02 region
# by ~1e-3 by itself, so its bound is 4e-3.
STORAGE_BOUND = {mexct.F16: 1e-3, mexct.BF16: 4e-3}


@pytest.mark.parametrize("storage", [mexct.F16, mexct.BF16])
def test_storage_modes(gpu, storage):
    from mexct import phantoms
    W, H, n = 64, 48, 16
    vol = phantoms.ellipsoid_phantom(W, H, 1, 5, 24.0, 0.02, (3.0, 10.0), (0.005, 0.03), 20.0, 15.0)
    ang = _angles(n)
    fp = mexct.CUDAmex_FP(vol, 1000.0, ang, n, 2, storage=storage)
    fp64 = O.fp(vol, 1000.0, ang, n, 2, f64=True)
    e = np.linalg.norm(fp - fp64) / np.linalg.norm(fp64)
    print(f"FP storage={storage} relRMSE vs FP64 = {e:.3e}")
    assert e <= STORAGE_BOUND[storage]
    bp = mexct.CUDAmex_BP(np.asfortranarray(fp64.astype(np.float32)), 1000.0, ang, 2, storage=storage)
    bp64 = O.bp(np.asfortranarray(fp64.astype(np.float32)), 1000.0, ang, 2, f64=True)
    e = np.linalg.norm(bp - bp64) / np.linalg.norm(bp64)
    print(f"BP storage={storage} relRMSE vs FP64 = {e:.3e}")
    assert e <= STORAGE_BOUND[storage]


# OSC-TV of the few-view cone workload (12 ellipsoids in a cylinder, Poisson
# transmission at I0=1e5, 10 subsets, 20 iterations, 10 TV steps, first-pick
# seed 303) at 64^3 with 60 views, against the FP64 oracle.  FP16/BF16
# storage keeps the projected volume copy and the backprojected ratio terms
# in that type; the bounds are those of the single FP and BP.
_OSC64 = {}


def _osc_case_64():
    if not _OSC64:
        from mexct import phantoms
        W, n = 64, 60
        vol = phantoms.osc_phantom(W, seed=3)
        ang = phantoms.angles_uniform(n, 2 * np.pi)
        fp = O.fp(vol, 1000.0, ang, n, 2)
        pre, post = phantoms.poisson_transmission(fp, 1e5, seed=33)
        picks = mexct.first_picks(303, 10, 10, 1.0, 20)
        ref = O.osc_tv_recon(pre, post, ang, 1000.0, 2, 20, 0.1, 10, 10, 1.0, 10, picks, f64=True)
        _OSC64.update(pre=pre, post=post, ang=ang, picks=picks, ref=ref)
    return _OSC64


@pytest.mark.parametrize("storage", [mexct.F32, mexct.F16, mexct.BF16])
def test_osc_tv_storage_vs_fp64(gpu, storage):
    c = _osc_case_64()
    got = mexct.OSC_TV_recon(c["pre"], c["post"], c["ang"], {"type": "cone3d", "SAD": 1000.0}, 20,
                             0.1, 10, 10, 1.0, tv_steps=10, picks=c["picks"], storage=storage)
    ref = c["ref"]
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"OSC-TV storage={storage} relRMSE vs FP64 = {e:.3e}")
    assert e <= {mexct.F32: 1e-4}.get(storage, STORAGE_BOUND.get(storage))


@pytest.mark.parametrize("storage", [mexct.F16, mexct.BF16])
def test_osc_iter_storage_vs_fp64(gpu, storage):
    from mexct import phantoms
    W, H, n = 48, 32, 24
    vol = phantoms.osc_phantom(W, seed=5)[:, :, :H].copy(order="F")
    ang = _angles(n)
    fp = O.fp(vol, 500.0, ang, n, 2)
    pre, post = phantoms.poisson_transmission(fp, 1e5, seed=7)
    mu0 = np.asfortranarray(np.full((W, W, H), 0.015, np.float32))
    picks = np.array([2, 0, 3, 1], np.int32)
    mu, pr, po = mu0.copy(order="F"), pre.copy(order="F"), post.copy(order="F")
    mexct.CUDAmex_oscIter(mu, pr, po, 4, picks, ang, 500.0, 2, storage=storage)
    mu64, _, _ = O.osc_iter(mu0, pre, post, 4, picks, ang, 500.0, 2, f64=True)
    e = np.linalg.norm(mu - mu64) / np.linalg.norm(mu64)
    print(f"oscIter storage={storage} relRMSE vs FP64 = {e:.3e}")
    assert e <= STORAGE_BOUND[storage]


def _round_storage(a, storage):
    """The values the device stores for `a` in FP16 / BF16 (round to nearest even)."""
    a = np.asarray(a, dtype=np.float32)
    if storage == mexct.F16:
        r = a.astype(np.float16).astype(np.float32)
    else:
        u = a.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        r = u.astype(np.uint32).view(np.float32)
    return np.asfortranarray(r)


# The storage forms change what is stored; FP and BP with 2-byte storage
# also step their sample coordinates in texel space (FP) or from z0 directly
# (BP), whose positions differ from the reference's float rounding by a few
# ulps.  Against the oracle run on the rounded operand they match to 1e-4
# relative max and 2e-5 relative L2 (random data, the steepest gradients);
# a wrong tap, weight or half-select in the mixed-precision lerps is off by
# orders of magnitude more.  The F16 projectors' first lerp level (along the
# detector columns in the BP, along x in the FP) runs in packed f16
# arithmetic, two roundings of at most 2^-11 of the tap values each: they
# are held to 1e-3 relative max and 2e-4 relative L2, still far below what a
# wrong tap or half-select gives (> 1e-2).
F16_LERP_BOUND = (1e-3, 2e-4)
@pytest.mark.parametrize("storage", [mexct.F16, mexct.BF16])
@pytest.mark.parametrize("geom", GEOMS)
def test_fp_storage_on_rounded_volume(gpu, geom, storage):
    W, H, n = 40, 36, 8
    vol = _vol(W, H, 61 + geom)
    ang = _angles(n)
    got = mexct.CUDAmex_FP(vol, 400.0, ang, n, geom, storage=storage)
    ref = O.fp(_round_storage(vol, storage), 400.0, ang, n, geom)
    tol = F16_LERP_BOUND if storage == mexct.F16 else (1e-4, 2e-5)
    _check(got, ref, f"FP storage={storage} geom={geom} vs rounded", *tol)


# SAD 400 takes the row-quad sinogram (one load per two z-voxels); at SAD 30
# the cone magnification reaches 2.7 and the row-paired one is used.
@pytest.mark.parametrize("storage", [mexct.F16, mexct.BF16])
@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("sad", [400.0, 30.0])
def test_bp_storage_on_rounded_sinogram(gpu, geom, storage, sad):
    W, H, n = 40, 36, 8
    P = np.asfortranarray(_rng(71 + geom).uniform(0, 1, (W, H, n)).astype(np.float32))
    ang = _angles(n)
    got = mexct.CUDAmex_BP(P, sad, ang, geom, storage=storage)
    ref = O.bp(_round_storage(P, storage), sad, ang, geom)
    tol = F16_LERP_BOUND if storage == mexct.F16 else (1e-4, 2e-5)
    _check(got, ref, f"BP storage={storage} geom={geom} sad={sad} vs rounded", *tol)


# The 2-byte cone row-quad BP at its limits: magnification up to 1.98 (SAD
# just above 2.02 hw), heights that are not a multiple of the z-range (the
# last range loads rows past H, which the row clamp maps to the edge) and
# detectors much taller than wide.
@pytest.mark.parametrize("W,H,n,sad", [(48, 100, 10, 47.0), (64, 70, 12, 90.0),
                                       (40, 150, 6, 400.0), (96, 33, 9, 1000.0)])
def test_bp_cone_row_quad_limits(gpu, W, H, n, sad):
    P = np.asfortranarray(_rng(W + H).uniform(0, 1, (W, H, n)).astype(np.float32))
    ang = _angles(n)
    got = mexct.CUDAmex_BP(P, sad, ang, 2, storage=mexct.F16)
    Pr = _round_storage(P, mexct.F16)
    # against the FP64 twin: the FP32 oracle replays the reference's running
    # row sum, whose drift over tall detectors (H = 150: ~1e-4 relative) the
    # 2-byte path's direct row start does not share
    ref = O.bp(Pr, sad, ang, 2, f64=True)
    m32, _ = rel_errors(got, O.bp(Pr, sad, ang, 2))
    print(f"  (vs the FP32 oracle: rel_max={m32:.3e})")
    _check(got, ref, f"BP f16 cone W={W} H={H} sad={sad} vs rounded, FP64", *F16_LERP_BOUND)


@pytest.mark.parametrize("storage", [mexct.F16, mexct.BF16])
def test_g3d_fp_storage_on_rounded_volume(gpu, storage):
    side, nsl, n = 24, 20, 4
    vol = _vol(side, nsl, 35)
    pts = np.asfortranarray(np.stack([_random_pts(10, 14, 60 + k) for k in range(n)], axis=3))
    ang = _angles(n)
    got = mexct.CUDAmex_general3D_FP(vol, ang, pts, n, storage=storage)
    ref = O.g3d_fp(_round_storage(vol, storage), ang, pts, n)
    tol = F16_LERP_BOUND if storage == mexct.F16 else (1e-4, 2e-5)
    _check(got, ref, f"g3D FP storage={storage} vs rounded", *tol)
