"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/mexct.h declares, and rejects bad arguments before touching
the device (the reference's argument errors, CUDAmex_FP.cu:352-353,463)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mexct
from mexct import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mexct.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(mexct_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ["mexct_fp", "mexct_bp", "mexct_g3d_fp", "mexct_g3d_bp", "mexct_osc_iter",
                 "mexct_tvmin", "mexct_prefilter", "mexct_osc_tv_recon"]:
        assert must in names and must + "_dev" in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python signature table covers them all
    assert set(_declared()) <= set(L.SIGNATURES)


def test_version():
    assert L.lib().mexct_version() >= 10000


def _f32(shape):
    return np.zeros(shape, np.float32, order="F")


def test_invalid_geometry():
    with pytest.raises(mexct.MexctError) as e:
        mexct.CUDAmex_FP(_f32((8, 8, 4)), 100.0, np.zeros(2, np.float32), 2, 3)
    assert e.value.code == L.EGEOM and "invalid geometry" in e.value.msg
    with pytest.raises(mexct.MexctError) as e:
        mexct.CUDAmex_BP(_f32((8, 4, 2)), 100.0, np.zeros(2, np.float32), -1)
    assert e.value.code == L.EGEOM


def test_nproj_exceeds_angles():
    with pytest.raises(mexct.MexctError) as e:
        mexct.CUDAmex_FP(_f32((8, 8, 4)), 100.0, np.zeros(2, np.float32), 3, 0)
    assert e.value.code == L.EINVAL
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_BP(_f32((8, 4, 5)), 100.0, np.zeros(4, np.float32), 0)


def test_non_single_inputs_rejected():
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_FP(np.zeros((8, 8, 4)), 100.0, np.zeros(2, np.float32), 2, 0)
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_FP(_f32((8, 8, 4)), 100.0, np.zeros(2), 2, 0)
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_FP(_f32((8, 6, 4)), 100.0, np.zeros(2, np.float32), 2, 0)


def test_osc_bad_subsets_and_picks():
    mu, pre, post = _f32((8, 8, 4)), _f32((8, 4, 6)), _f32((8, 4, 6))
    ang = np.zeros(6, np.float32)
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_oscIter(mu, pre, post, 0, np.zeros(1, np.int32), ang, 100.0, 2)
    with pytest.raises(mexct.MexctError):
        mexct.CUDAmex_oscIter(mu, pre, post, 3, np.zeros(2, np.int32), ang, 100.0, 2)
    with pytest.raises(mexct.MexctError):  # in-place args must be Fortran float32
        mexct.CUDAmex_oscIter(np.zeros((8, 8, 4), np.float32), pre, post, 1,
                              np.zeros(1, np.int32), ang, 100.0, 2)


def test_null_pointer_rejected():
    lib = L.lib()
    rc = lib.mexct_fp(None, 8, 8, C.c_float(1.0), None, 0, 0, 0, None, None)
    assert rc == L.EINVAL
    assert b"null" in lib.mexct_last_error()


def test_prefilter_needs_two_views():
    with pytest.raises(mexct.MexctError):
        mexct.preFilterSinogram(_f32((8, 4, 1)), {"type": "cone3d", "SAD": 100.0},
                                np.zeros(1, np.float32))


def test_parker_and_fbp_argument_checks():
    # rejected on the host before any device work
    sino = _f32((8, 4, 3))
    ang = np.zeros(3, np.float32)
    g = {"type": "fan3d", "SAD": 100.0}
    with pytest.raises(mexct.MexctError, match="q must be"):
        mexct.parker_weight(sino, ang, g, -1.0)
    with pytest.raises(mexct.MexctError, match="SAD"):
        mexct.parker_weight(sino, ang, {"type": "fan3d", "SAD": 0.0}, 1.0)
    with pytest.raises(mexct.MexctError, match="nProj exceeds"):
        mexct.parker_weight(sino, ang[:2], g, 1.0)
    with pytest.raises(mexct.MexctError, match="same shape"):
        mexct.FBP_recon(sino, _f32((8, 4, 2)), ang, g)
    with pytest.raises(mexct.MexctError, match="window"):
        L.check(L.lib().mexct_fbp_recon(sino.ctypes.data, sino.ctypes.data, 8, 4, 3,
                                        ang.ctypes.data, 3, 1, 100.0, 9, 1.0, -1.0,
                                        sino.ctypes.data, None))


def test_subset_helpers_match_oracle():
    import oracle as O

    for (s0, s1, p, K) in [(64, 2, 0.2, 10), (10, 10, 1.0, 20), (32, 4, 0.5, 7)]:
        for k in range(1, K + 1):
            assert mexct.calc_num_subsets(s0, s1, p, K, k) == O.calc_num_subsets(s0, s1, p, K, k)
    for n in range(1, 40):
        for first in range(1, n + 1):
            assert np.array_equal(mexct.calc_subset_order(n, first), O.calc_subset_order(n, first))


def test_request_struct_layout_matches_header(tmp_path):
    """mexct_request as ctypes sees it must match the C compiler's layout."""
    import shutil
    import subprocess

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    fields = [f for f, _ in L.Request._fields_]
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "mexct.h"\nint main(void){'
                   'printf("%zu", sizeof(mexct_request));'
                   + "".join(f'printf(" %zu", offsetof(mexct_request, {f}));' for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got[0] == C.sizeof(L.Request)
    assert got[1:] == [getattr(L.Request, f).offset for f in fields]


def test_server_create_rejects_bad_arguments():
    h = C.c_void_p()
    assert L.lib().mexct_server_create(C.byref(h), 0, 1 << 30, 0, 1) == L.EINVAL
    assert L.lib().mexct_server_create(C.byref(h), 0, 0, 2, 1) == L.EINVAL
    assert L.lib().mexct_server_create(None, 0, 1 << 30, 2, 1) == L.EINVAL
    assert L.lib().mexct_server_submit(None, None, None) == L.EINVAL
    assert L.lib().mexct_server_destroy(None) == L.OK


def test_prefilter_noalloc_rejects_tables_beyond_scratch_before_the_device():
    # the size check runs before the context is looked at, so it needs no GPU
    rc = L.lib().mexct_prefilter_noalloc_dev(None, None, 2048, 2048, 10, 2, 1000.0, None, 3, 1.0,
                                             -1.0, None)
    assert rc == L.EINVAL
    msg = L.lib().mexct_last_error().decode()
    assert "allocation-free" in msg and str(16 << 20) in msg
    rc = L.lib().mexct_prefilter_noalloc_dev(None, None, 64, 64, 10, 2, 1000.0, None, 3, 1.0,
                                             -1.0, None)
    assert rc == L.EINVAL and "ctx is null" in L.lib().mexct_last_error().decode()


def test_device_wrappers_check_sizes_before_the_library():
    # The _dev entry points take device pointers without lengths, so the
    # Context methods check every size first; these raise before any
    # library call (the context below has no handle and no GPU behind it).
    torch = pytest.importorskip("torch")
    from mexct import device

    ctx = object.__new__(device.Context)
    ctx._h, ctx.order_with_torch, ctx.stream = None, False, None
    z = lambda *s: torch.zeros(s)  # noqa: E731
    ang4 = np.zeros(4, np.float32)
    bad = [
        lambda: ctx.g3d_fp(z(8, 8, 8), ang4, z(3, 6, 5, 7), 4, z(4, 5, 7)),   # 3 views of pts
        lambda: ctx.g3d_fp(z(8, 8, 8), np.zeros(2), z(6, 5, 7), 4, z(4, 5, 7)),  # 2 angles
        lambda: ctx.g3d_fp(z(8, 8, 8), ang4, z(6, 5, 7), 4, z(4, 5, 6)),      # out too small
        lambda: ctx.g3d_bp(z(4, 5, 7), ang4, z(2, 6, 5, 7), 8, 8, z(8, 8, 8)),
        lambda: ctx.g3d_bp(z(4, 5, 7), ang4, z(6, 5, 7), 8, 8, z(8, 8, 7)),
        lambda: ctx.osc_iter(z(4, 8, 8), z(6, 4, 8), z(6, 4, 8), 3, [0, 1], np.zeros(6), 100, 2),
        lambda: ctx.osc_iter(z(4, 8, 8), z(6, 4, 8), z(6, 4, 8), 2, [0, 1], np.zeros(5), 100, 2),
        lambda: ctx.osc_iter(z(4, 8, 8), z(6, 4, 8), z(5, 4, 8), 2, [0, 1], np.zeros(6), 100, 2),
        lambda: ctx.prefilter(z(6, 4, 8), 2, 100, np.zeros(5), 3, 1.0, z(6, 4, 8)),
        lambda: ctx.prefilter(z(6, 4, 8), 2, 100, np.zeros(6), 3, 1.0, z(6, 4, 7)),
        lambda: ctx.parker_weight(z(6, 4, 8), np.zeros(6), 2, 100, 1.0, z(5, 4, 8)),
        lambda: ctx.fbp_recon(z(6, 4, 8), z(6, 4, 8), np.zeros(6), 2, 100, 3, 1.0, z(4, 8, 7)),
        lambda: ctx.osc_tv_recon(z(6, 4, 8), z(6, 4, 8), np.zeros(6), 100, 2, 5, 0.1, 4, 2, 1.0,
                                 [1, 1, 1], z(4, 8, 8)),
        lambda: ctx.osc_tv_recon(z(6, 4, 8), z(6, 4, 8), np.zeros(6), 100, 2, 3, 0.1, 4, 2, 1.0,
                                 [1, 1, 1], z(4, 8, 9)),
        lambda: ctx.tvmin(z(8, 8), 0.1),
    ]
    for i, f in enumerate(bad):
        with pytest.raises(mexct.MexctError) as e:
            f()
        assert e.value.code == L.EINVAL, i
        assert "device tensor" not in e.value.msg, (i, e.value.msg)  # a size check, not _dptr


def test_server_footprint_counts_the_abi_workspace():
    # The reservation of a request covers what its entry point allocates:
    # the FP's x-pair volume copy (about 2x the haloed volume) and the cone
    # f32 BP's row-start chain (up to 1 GiB), computed from the ABI's own
    # layouts.  No GPU is involved.
    from mexct import serve

    W = H = 256
    n = 360
    v, s = W * W * H * 4, W * H * n * 4
    dual = (W + 3) * (W + 4) * (H + 4) * 2 * 4
    fp = serve.request_footprint(L.OP_FP, W, H, n, geom=2)
    assert fp >= v + s + dual
    bp_cone = serve.request_footprint(L.OP_BP, W, H, n, geom=2)
    bp_par = serve.request_footprint(L.OP_BP, W, H, n, geom=0)
    assert bp_cone - bp_par > 500 << 20  # the chain
    assert serve.request_footprint(L.OP_FDK, W, H, n, geom=2) > bp_cone
    # OSC-TV: the fewer subsets, the larger the per-subset ratio copies
    a = serve.request_footprint(L.OP_OSCTV, W, H, n, osc_iters=5, s_init=10, s_final=10)
    b = serve.request_footprint(L.OP_OSCTV, W, H, n, osc_iters=5, s_init=10, s_final=1)
    assert b > a > 6 * s
    with pytest.raises(mexct.MexctError):
        serve.request_footprint(9, W, H, n)


def test_osc_tv_iteration_range_checked_before_the_device():
    f = np.zeros(4, np.float32)
    p = np.ones(5, np.int32)
    for k0, k1 in ((0, 2), (3, 2), (2, 6)):
        rc = L.lib().mexct_osc_tv_iterations(f.ctypes.data, f.ctypes.data, 2, 1, 2, f.ctypes.data,
                                             4, 100.0, 2, 5, 0.1, 2, 2, 1.0, p.ctypes.data, k0, k1,
                                             f.ctypes.data, None)
        assert rc == L.EINVAL and "iterations" in L.lib().mexct_last_error().decode()
    with pytest.raises(mexct.MexctError):
        pre = np.ones((8, 4, 4), np.float32, order="F")
        mexct.OSC_TV_recon(pre, pre, np.zeros(4, np.float32), {"type": "cone3d", "SAD": 100.0},
                           5, 0.1, 2, 2, 1.0, picks=np.ones(5, np.int32), iterations=(2, 3))
