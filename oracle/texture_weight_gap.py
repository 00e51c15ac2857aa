"""How far the exact-weight oracle sits from NVIDIA's texture unit.

TEST INFRASTRUCTURE (a measurement over the oracle only; no GPU).  The
reference fetches through CUDA's linear filter, whose fraction is held in
9-bit fixed point with 8 fraction bits (CUDA Programming Guide, texture
fetching appendix); this build and its oracle use exact FP32 fractions.
This script runs the FP32 oracle both ways on the two standard-geometry
workloads and prints the gap, relative max (max|a-b| / max|b|) and
relative L2 (||a-b|| / ||b||):

  parallel 128^3 x 180 (cylinder r=48 + 5 ellipsoids, seed 1): full FP,
      then full BP of that FP;
  cone 384^3 x 512 (cylinder r=140, 0.02, SAD 1000): FP of 4096 seeded
      rays, BP of 64 seeded in-circle voxel columns of the full FP of a
      64-view slice of the workload (all 512 angles kept for the FP rays).

    python oracle/texture_weight_gap.py [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mex-cuda-ct_amd"))

import oracle as O  # noqa: E402
from mexct import phantoms as ph  # noqa: E402


def gap(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = a - b
    return {"rel_max": float(np.abs(d).max() / np.abs(b).max()),
            "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(b))}


def both(fn):
    exact = fn()
    with O.texture_weight_bits(8):
        hw = fn()
    return gap(hw, exact)


def parallel_128():
    vol = ph.par_phantom_128(1)
    ang = ph.angles_uniform(180, np.pi)
    out = {"fp": both(lambda: O.fp(vol, 1000.0, ang, 180, 0))}
    sino = O.fp(vol, 1000.0, ang, 180, 0)
    out["bp"] = both(lambda: O.bp(sino, 1000.0, ang, 0))
    return out


def cone_384(rays=4096, cols=64, views=64):
    W = H = 384
    n = 512
    vol = ph.cone_cylinder_384()
    ang = ph.angles_linspace(n, 2 * np.pi)
    r = np.random.Generator(np.random.PCG64(8))
    ix = r.integers(0, W, rays)
    iy = r.integers(0, H, rays)
    vw = r.integers(0, n, rays)
    out = {"fp_rays": both(lambda: O.fp_rays(vol, 1000.0, ang, 2, ix, iy, vw))}
    sub = ang[:: n // views]
    sino = O.fp(vol, 1000.0, sub, sub.size, 2)
    hw = W // 2 - 1
    cx, cy = [], []
    while len(cx) < cols:
        x, y = r.integers(0, W, 2)
        if (x - hw) ** 2 + (y - hw) ** 2 < hw * hw:
            cx.append(x)
            cy.append(y)
    out["bp_cols"] = both(lambda: O.bp_cols(sino, 1000.0, sub, 2, np.array(cx), np.array(cy)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        O.set_threads(a.threads)
    t = time.time()
    res = {"parallel_128x180": parallel_128(), "cone_384x512_sampled": cone_384()}
    res["seconds"] = round(time.time() - t, 1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
