"""mexct -- MI355X (gfx950) CT projection and iterative-reconstruction
operators with the MEX entry points of mex-cuda-CT.

Host-array mirror of the MEX calls (numpy, MATLAB layout):
    CUDAmex_FP, CUDAmex_BP, CUDAmex_general3D_FP, CUDAmex_general3D_BP,
    CUDAmex_oscIter, CUDAmex_TVmin_3D,
and of the MATLAB drivers around them:
    preFilterSinogram, parker_weight, OSC_TV_recon, FBP_recon.
Device-resident use: mexct.device.Context.
All compute runs in libmexct.so (HIP); there is no CPU fallback.
"""
from ._lib import BF16, F16, F32, MexctError, LIB_PATH, lib  # noqa: F401
from .ops import (  # noqa: F401
    CUDAmex_BP,
    CUDAmex_FP,
    CUDAmex_general3D_BP,
    CUDAmex_general3D_FP,
    CUDAmex_oscIter,
    CUDAmex_TVmin_3D,
    FBP_recon,
    GEOM,
    OSC_TV_recon,
    WINDOWS,
    calc_num_subsets,
    calc_subset_order,
    first_picks,
    parker_weight,
    preFilterSinogram,
    set_autotune,
    tuned_config,
)
