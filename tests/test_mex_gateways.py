"""The C++ MEX gateways (mexFunction over the C ABI), called as MATLAB
would through the in-tree mex.h shim.  Argument errors need no GPU; the
outputs are compared with the mexct host entry points on the GPU."""
import numpy as np
import pytest

import mexct
from mex_harness import MexCallError, call


def f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def test_fp_too_few_arguments():
    with pytest.raises(MexCallError) as e:
        call("CUDAmex_FP", f32(np.zeros((4, 4, 4))), 100.0)
    assert e.value.ident == "mex_conebeam_FP:invalidInput"


def test_fp_invalid_geometry_message():
    with pytest.raises(MexCallError) as e:
        call("CUDAmex_FP", f32(np.zeros((4, 4, 4))), 100.0, f32(np.zeros(2)), 2.0, 7.0)
    assert e.value.msg == "invalid geometry"


def test_bp_ids_and_class_checks():
    with pytest.raises(MexCallError) as e:
        call("CUDAmex_BP", f32(np.zeros((4, 4, 2))), 100.0)
    assert e.value.ident == "mex_conebeam_fdk:invalidInput"
    with pytest.raises(MexCallError):
        call("CUDAmex_BP", np.zeros((4, 4, 2)), 100.0, f32(np.zeros(2)), 0.0)


def test_tvmin_needs_exactly_two_args():
    with pytest.raises(MexCallError) as e:
        call("CUDAmex_TVmin_3D", f32(np.zeros((4, 4, 4))))
    assert "2 arguments expected" in e.value.msg


def test_osc_picks_must_be_int32():
    mu, pre, post = f32(np.zeros((8, 8, 4))), f32(np.ones((8, 4, 4))), f32(np.ones((8, 4, 4)))
    with pytest.raises(MexCallError):
        call("CUDAmex_oscIter", mu, pre, post, 2.0, np.zeros(2), f32(np.zeros(4)), 100.0, 2.0,
             nlhs=0)


def test_g3d_bp_per_view_pts_must_cover_every_view():
    # [W,H,6,k] pts with k < nProj would be read past its end
    W, H, n = 4, 3, 5
    P = f32(np.zeros((W, H, n)))
    pts = f32(np.zeros((W, H, 6, n - 1)))
    with pytest.raises(MexCallError) as e:
        call("CUDAmex_general3D_BP", P, f32(np.zeros(n)), pts, 4.0, 4.0)
    assert "fewer views" in e.value.msg
    with pytest.raises(mexct.MexctError) as e2:
        mexct.CUDAmex_general3D_BP(P, f32(np.zeros(n)), pts, 4, 4)
    assert e2.value.code == mexct._lib.EINVAL
    with pytest.raises(MexCallError):
        call("CUDAmex_general3D_FP", f32(np.zeros((4, 4, 4))), f32(np.zeros(n)), pts, float(n))


@pytest.mark.gpu
def test_gateways_match_host_entry_points(gpu):
    r = np.random.Generator(np.random.PCG64(3))
    W, H, n = 24, 16, 6
    vol = f32(r.uniform(size=(W, W, H)))
    ang = f32(np.arange(n) * 0.5)
    for geom in (0, 1, 2):
        (fp,) = call("CUDAmex_FP", vol, 300.0, ang, float(n), float(geom))
        np.testing.assert_array_equal(fp, mexct.CUDAmex_FP(vol, 300.0, ang, n, geom))
        (bp,) = call("CUDAmex_BP", fp, 300.0, ang, float(geom))
        np.testing.assert_array_equal(bp, mexct.CUDAmex_BP(fp, 300.0, ang, geom))
    # 4-argument CUDAmex_FP defaults to parallel beam
    (fp4,) = call("CUDAmex_FP", vol, 300.0, ang, float(n))
    np.testing.assert_array_equal(fp4, mexct.CUDAmex_FP(vol, 300.0, ang, n, 0))
    from mexct import phantoms
    pts = phantoms.parallel_pts(W, H)
    (g,) = call("CUDAmex_general3D_FP", vol, ang, pts, float(n))
    np.testing.assert_array_equal(g, mexct.CUDAmex_general3D_FP(vol, ang, pts, n))
    (gb,) = call("CUDAmex_general3D_BP", g, ang, pts, float(W), float(H))
    ref = mexct.CUDAmex_general3D_BP(g, ang, pts, W, H)
    np.testing.assert_allclose(gb, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
    (tv,) = call("CUDAmex_TVmin_3D", vol, 0.01)
    np.testing.assert_array_equal(tv, mexct.CUDAmex_TVmin_3D(vol, 0.01))
    # in-place OSC: the MATLAB arrays come back mutated
    mu = f32(r.uniform(0, 0.02, size=(W, W, H)))
    pre = f32(r.uniform(5e4, 1e5, size=(W, H, n)))
    post = f32(pre * 0.5)
    mu2, pre2, post2 = mu.copy(order="F"), pre.copy(order="F"), post.copy(order="F")
    picks = np.array([1, 0], np.int32)
    call("CUDAmex_oscIter", mu, pre, post, 2.0, picks, ang, 300.0, 2.0, nlhs=0,
         inplace=(0, 1, 2))
    mexct.CUDAmex_oscIter(mu2, pre2, post2, 2, picks, ang, 300.0, 2)
    np.testing.assert_array_equal(mu, mu2)
    np.testing.assert_array_equal(pre, pre2)
    np.testing.assert_array_equal(post, post2)
