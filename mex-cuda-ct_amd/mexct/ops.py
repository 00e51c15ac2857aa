"""The six MEX entry points of mex-cuda-CT, mirrored over libmexct's C ABI.

Each function takes and returns numpy arrays in MATLAB's layout: float32,
column-major (Fortran order), volume dims (W, W, H), sinogram dims
(W, H, nProj).  Argument meaning and order follow the reference's MEX
gateways; the cited lines are the reference's.  Arrays that the reference
would reinterpret as `single` (mxGetData cast to float*) must already be
float32 here: anything else raises instead of being read as garbage.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import F32, MexctError, EINVAL, check, lib, opts

GEOM = {"par3d": 0, "fan3d": 1, "cone3d": 2, "cone": 2}
WINDOWS = {"ramp": 0, "shepp-logan": 1, "cosine": 2, "hamming": 3, "hann": 4, "none": 5}


def _single(a, name, writable=False):
    """A float32 Fortran-ordered array (copy-free when it already is one)."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float32:
        raise MexctError(EINVAL, f"{name} must be a float32 (single) array")
    if writable:
        if not a.flags.f_contiguous:
            raise MexctError(EINVAL, f"{name} is updated in place and must be Fortran-contiguous")
        return a
    return np.asfortranarray(a)


def _dims3(a):
    s = a.shape + (1,) * (3 - a.ndim)
    return s[0], s[1], s[2]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _vec(a, dtype, name):
    v = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
    if v.dtype != dtype:
        raise MexctError(EINVAL, f"{name} must be {np.dtype(dtype).name}")
    return v


def CUDAmex_FP(vol, SAD, angles, nProjections, geomFlag, storage=F32, long_ray=0):
    """fp = CUDAmex_FP(vol, SAD, angles, nProj, geomFlag) (CUDAmex_FP.cu:350-477).

    vol [W, W, H] -> fp [W, H, nProj].  SAD goes through float, nProj and
    geomFlag through int, as mxGetScalar -> C conversion in the reference.
    long_ray: F16/BF16 storage sums rays of more samples with compensated
    summation (0: the default 1024, negative: never)."""
    vol = _single(vol, "vol")
    W, Wy, H = _dims3(vol)
    if Wy != W:
        raise MexctError(EINVAL, f"volume must be [W,W,H], got {vol.shape}")
    ang = _vec(angles, np.float32, "angles")
    nproj = int(nProjections)
    out = np.zeros((W, H, max(nproj, 0)), dtype=np.float32, order="F")
    check(lib().mexct_fp(_ptr(vol), W, H, float(np.float32(SAD)), _ptr(ang), ang.size, nproj,
                         int(geomFlag), _ptr(out), opts(storage, long_ray=int(long_ray))))
    return out


def CUDAmex_BP(P, SAD, angles, geomFlag, storage=F32, zrange=0):
    """bp = CUDAmex_BP(P, SAD, angles, geomFlag) (CUDAmex_BP.cu:223-324).

    P [W, H, nProj] -> bp [W, W, H]."""
    P = _single(P, "projections")
    W, H, n = _dims3(P)
    ang = _vec(angles, np.float32, "angles")
    out = np.zeros((W, W, H), dtype=np.float32, order="F")
    check(lib().mexct_bp(_ptr(P), W, H, n, float(np.float32(SAD)), _ptr(ang), ang.size,
                         int(geomFlag), _ptr(out), opts(storage, 0, zrange)))
    return out


def CUDAmex_general3D_FP(vol, angles, pts, nProjections, storage=F32, long_ray=0):
    """fp = CUDAmex_general3D_FP(vol, angles, pts, nProj) (CUDAmex_general3D_FP.cu:154-252).

    pts [W, H, 6] is shared by every view and rotated by its angle; a
    [W, H, 6, nProj] array gives each view its own rays (additive form)."""
    vol = _single(vol, "vol")
    pts = _single(pts, "pts")
    side, _, nsl = _dims3(vol)
    if pts.ndim not in (3, 4) or pts.shape[2] != 6:
        raise MexctError(EINVAL, f"pts must be [W,H,6] or [W,H,6,nProj], got {pts.shape}")
    ppv = 1 if pts.ndim == 4 else 0
    W, H = pts.shape[0], pts.shape[1]
    nproj = int(nProjections)
    if ppv and pts.shape[3] < nproj:
        raise MexctError(EINVAL, "per-view pts has fewer views than nProj")
    ang = _vec(angles, np.float32, "angles")
    out = np.zeros((W, H, max(nproj, 0)), dtype=np.float32, order="F")
    check(lib().mexct_g3d_fp(_ptr(vol), side, nsl, _ptr(ang), ang.size, _ptr(pts), W, H, nproj,
                             ppv, _ptr(out), opts(storage, long_ray=int(long_ray))))
    return out


def CUDAmex_general3D_BP(P, angles, pts, M, N, mode=0):
    """bp = CUDAmex_general3D_BP(P, angles, pts, M, N) (CUDAmex_general3D_BP.cu:194-282).

    mode: 0 voxel bricks (default), 1 the same with deterministic fixed-point
    sums (the same bits on every run), 2 the ray-driven float atomics."""
    P = _single(P, "projections")
    pts = _single(pts, "pts")
    W, H, n = _dims3(P)
    if pts.ndim not in (3, 4) or pts.shape[0] != W or pts.shape[1] != H or pts.shape[2] != 6:
        raise MexctError(EINVAL, f"pts must be [W,H,6(,nProj)] matching P, got {pts.shape}")
    ppv = 1 if pts.ndim == 4 else 0
    if ppv and pts.shape[3] < n:
        raise MexctError(EINVAL, "per-view pts has fewer views than the projections")
    ang = _vec(angles, np.float32, "angles")
    M, N = int(M), int(N)
    out = np.zeros((M, M, N), dtype=np.float32, order="F")
    check(lib().mexct_g3d_bp(_ptr(P), W, H, n, _ptr(ang), ang.size, _ptr(pts), ppv, M, N,
                             _ptr(out), opts(g3d_bp=int(mode))))
    return out


def CUDAmex_oscIter(mu, pre, post, numSubsets, subsetPicks, angles, SAD, geomFlag, storage=F32):
    """CUDAmex_oscIter(mu, pre, post, numSubsets, picks, angles, SAD, geomFlag).

    Updates mu, pre and post IN PLACE, as the reference mutates its MATLAB
    inputs (CUDAmex_oscIter.cu:527-528, :564-570).  picks are int32, 0-based.
    storage F16/BF16 keeps the projected volume copy and the backprojected
    ratio terms in that type; mu, pre and post stay float32."""
    mu = _single(mu, "mu", writable=True)
    pre = _single(pre, "pre", writable=True)
    post = _single(post, "post", writable=True)
    W, _, H = _dims3(mu)
    pw, ph, n = _dims3(pre)
    if (pw, ph) != (W, H) or post.shape != pre.shape:
        raise MexctError(EINVAL, "pre/post must be [W,H,nProj] for a [W,W,H] volume")
    picks = _vec(subsetPicks, np.int32, "subsetPicks")
    ang = _vec(angles, np.float32, "angles")
    check(lib().mexct_osc_iter(_ptr(mu), W, H, _ptr(pre), _ptr(post), n, int(numSubsets),
                               _ptr(picks), picks.size, _ptr(ang), ang.size,
                               float(np.float32(SAD)), int(geomFlag), opts(int(storage))))


def CUDAmex_TVmin_3D(vol, c, steps=20):
    """out = CUDAmex_TVmin_3D(vol, c) (CUDAmex_TVmin_3D.cu:162-246); `steps`
    defaults to the reference's hard-coded 20."""
    vol = _single(vol, "vol")
    Nx, Ny, Nz = _dims3(vol)
    out = np.zeros_like(vol, order="F")
    check(lib().mexct_tvmin(_ptr(vol), Nx, Ny, Nz, float(np.float32(c)), _ptr(out),
                            opts(F32, int(steps))))
    return out


def _parker_q(q):
    return -1.0 if q is None else float(q)


def _geom_sad(geom):
    g = GEOM[str(geom["type"]).lower()]
    return g, float(geom.get("SAD", 0.0))


def preFilterSinogram(sino, geom, angles, filter="hamming", frequencyCutoff=1.0, parker_q=None):
    """filtered = preFilterSinogram(sino, geom, angles, filter, cutoff[, parker_q])
    (preFilterSinogram.m:29-152).  geom is a dict with 'type' ('par3d' |
    'fan3d' | 'cone3d') and 'SAD'.  parker_q None is the 5-argument call
    (no Parker weighting)."""
    sino = _single(sino, "sino")
    W, H, n = _dims3(sino)
    g, sad = _geom_sad(geom)
    w = WINDOWS[str(filter).lower()]
    ang = _vec(angles, np.float32, "angles")
    out = np.zeros((W, H, n), dtype=np.float32, order="F")
    check(lib().mexct_prefilter(_ptr(sino), W, H, n, g, sad, _ptr(ang), ang.size, w,
                                float(frequencyCutoff), _parker_q(parker_q), _ptr(out)))
    return out


def parker_weight(sino, angles, geom, q_value):
    """weightedSino = parker_weight(sino, angles, geom, q_value) (parker_weight.m:1-79)."""
    sino = _single(sino, "sino")
    W, H, n = _dims3(sino)
    g, sad = _geom_sad(geom)
    ang = _vec(angles, np.float32, "angles")
    out = np.zeros((W, H, n), dtype=np.float32, order="F")
    check(lib().mexct_parker_weight(_ptr(sino), W, H, n, _ptr(ang), ang.size, g, sad,
                                    float(q_value), _ptr(out)))
    return out


def calc_num_subsets(s_init, s_final, p, n_iter, k):
    """calcNumSubsets (OSC_TV_recon.m:323-325)."""
    return lib().mexct_calc_num_subsets(int(s_init), int(s_final), float(p), int(n_iter), int(k))


def calc_subset_order(n, first_pick):
    """calcSubsetOrder (OSC_TV_recon.m:327-349) from a given first pick (1-based)."""
    out = np.zeros(int(n), dtype=np.int32)
    check(lib().mexct_calc_subset_order(int(n), int(first_pick), _ptr(out)))
    return out


def first_picks(seed, s_init, s_final, p, n_iter):
    """Seeded replacement of MATLAB's randperm(n, 1) draw per iteration."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.array([1 + rng.integers(calc_num_subsets(s_init, s_final, p, n_iter, k))
                     for k in range(1, n_iter + 1)], dtype=np.int32)


def OSC_TV_recon(pre, post, scanAngles, geom, numIter, TV_constant, numSubsInit, numSubsFinal,
                 pVal, tv_steps=20, seed=0, picks=None, storage=F32, iterations=None, mu=None):
    """recon = OSC_TV_recon(pre, post, angles, geom, nIter, TV, sInit, sFinal, p)
    (OSC_TV_recon.m:1-102) with all iterations on the device.

    iterations=(k_begin, k_end) runs only those iterations (1-based,
    inclusive) of the numIter-iteration schedule, from mu = the volume after
    iteration k_begin - 1 (not needed when k_begin is 1): a checkpointed run
    resumed this way gives the bits of the uninterrupted one."""
    pre = _single(pre, "pre")
    post = _single(post, "post")
    W, H, n = _dims3(pre)
    g = GEOM[str(geom["type"]).lower()]
    ang = _vec(scanAngles, np.float32, "angles")
    if picks is None:
        picks = first_picks(seed, numSubsInit, numSubsFinal, pVal, int(numIter))
    picks = _vec(picks, np.int32, "picks")
    if picks.size < int(numIter):
        raise MexctError(EINVAL, "picks needs numIter entries")
    if iterations is not None:
        k0, k1 = (int(k) for k in iterations)
        out = np.zeros((W, W, H), dtype=np.float32, order="F")
        if k0 > 1:
            if mu is None:
                raise MexctError(EINVAL, "resuming after iteration 1 needs the volume mu")
            out[...] = _single(mu, "mu")
        check(lib().mexct_osc_tv_iterations(
            _ptr(pre), _ptr(post), W, H, n, _ptr(ang), ang.size, float(np.float32(geom["SAD"])),
            g, int(numIter), float(np.float32(TV_constant)), int(numSubsInit), int(numSubsFinal),
            float(pVal), _ptr(picks), k0, k1, _ptr(out), opts(int(storage), int(tv_steps))))
        return out
    mu = np.zeros((W, W, H), dtype=np.float32, order="F")
    check(lib().mexct_osc_tv_recon(_ptr(pre), _ptr(post), W, H, n, _ptr(ang), ang.size,
                                   float(np.float32(geom["SAD"])), g, int(numIter),
                                   float(np.float32(TV_constant)), int(numSubsInit),
                                   int(numSubsFinal), float(pVal), _ptr(picks), _ptr(mu),
                                   opts(int(storage), int(tv_steps))))
    return mu


def FBP_recon(pre, post, scanAngles, geom, filter="hamming", filterCutoff=1.0, parker_q=None,
              storage=F32):
    """recon = FBP_recon(pre, post, angles, geom, filter, cutoff[, parker_q])
    (FBP_recon.m:1-40): -log(post./pre) with NaN/Inf -> 0, the pre-filter
    and the BP of the geometry, all on the device."""
    pre = _single(pre, "pre")
    post = _single(post, "post")
    W, H, n = _dims3(pre)
    if post.shape != pre.shape:
        raise MexctError(EINVAL, "pre and post must have the same shape")
    g, sad = _geom_sad(geom)
    w = WINDOWS[str(filter).lower()]
    ang = _vec(scanAngles, np.float32, "angles")
    vol = np.zeros((W, W, H), dtype=np.float32, order="F")
    check(lib().mexct_fbp_recon(_ptr(pre), _ptr(post), W, H, n, _ptr(ang), ang.size, g, sad, w,
                                float(filterCutoff), _parker_q(parker_q), _ptr(vol),
                                opts(int(storage))))
    return vol


def set_autotune(mode):
    """Block-order autotuning of the FP/BP kernels for the process: 1 tunes
    each problem shape at its first call (timing every candidate on that
    call's data), 2 at its second, 0 off."""
    check(lib().mexct_set_autotune(int(mode)))


def tuned_config(op, geom, storage, W, H, nProj):
    """(found, FP views-per-chunk log2 or BP super-tile side, BP z-range,
    winner's time in us) for a problem shape; op 0 FP, 1 BP."""
    cfg = np.zeros(4, dtype=np.int32)
    check(lib().mexct_tuned_config(int(op), int(geom), int(storage), int(W), int(H), int(nProj),
                                   _ptr(cfg)))
    return tuple(int(x) for x in cfg)
