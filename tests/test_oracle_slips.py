"""The oracle's pins catch the mistakes they are meant to catch.

For each planted slip, a copy of oracle/oracle.cpp with that one mistake is
compiled with the oracle's own flags, and the pins that guard against it are
run on the copy: each must fail.  An unmodified copy built the same way must
pass them all, so a failure comes from the slip and not from the harness.
CPU only; the copies live in pytest's temporary directory.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

import oracle as O
import test_oracle_pins as P

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math",
            "-shared"]

# name -> (list of (text in oracle.cpp, replacement), pins that must fail)
SLIPS = {
    # OSC projection marched at CUDAmex_FP's cone step 1.0 instead of 0.5
    "osc_cone_step_1": (
        [("R(1) / R(2));\n          step = R(1) / R(2);", "R(1));")],
        [(P.test_osc_projection_is_border_step_half_march, 2),
         (P.test_osc_cone_axis_ray_closed_form,)]),
    # OSC projection with clamp addressing (CUDAmex_FP's) instead of border
    "osc_fp_clamp": (
        [("step, ORC_BORDER, nullptr);", "step, ORC_CLAMP, nullptr);")],
        [(P.test_osc_projection_is_border_step_half_march, 0),
         (P.test_osc_projection_is_border_step_half_march, 1),
         (P.test_osc_cone_axis_ray_closed_form,)]),
    # OSC dual backprojection with clamp addressing instead of border
    "osc_bp_clamp": (
        [("tex2<R, R>(tpre.data(), W, H, py, pz, ORC_BORDER)",
          "tex2<R, R>(tpre.data(), W, H, py, pz, ORC_CLAMP)"),
         ("tex2<R, R>(tpost.data(), W, H, py, pz, ORC_BORDER)",
          "tex2<R, R>(tpost.data(), W, H, py, pz, ORC_CLAMP)")],
        [(P.test_osc_update_equals_bp_of_ratio_sinograms, 1),
         (P.test_osc_update_equals_bp_of_ratio_sinograms, 2)]),
    # OSC fan stepping detector rows by UU (the cone rule) instead of 1
    "osc_fan_row_step_uu": (
        [("U = R(1);\n          UU = SR / (SR - s);", "U = SR / (SR - s);\n          UU = SR / (SR - s);")],
        [(P.test_osc_update_equals_bp_of_ratio_sinograms, 1)]),
    # subsets as contiguous view blocks instead of every numSubsets-th view
    "osc_subset_blocks": (
        [("for (int v = sel; v < nProj; v += numSubsets) views.push_back(v);",
          "for (int v = sel * (nProj / numSubsets); v < (sel + 1) * (nProj / numSubsets); ++v) "
          "views.push_back(v);")],
        [(P.test_osc_subsets_compose_in_pick_order, 0)]),
    # TV: the two xy-diagonal neighbours exchanged
    "tv_diag_swap": (
        [("R xmypz = in[c - sx + sy], xpymz = in[c + sx - sy];",
          "R xmypz = in[c + sx - sy], xpymz = in[c - sx + sy];")],
        [(P.test_tv_gradient_is_gradient_of_tv_functional,)]),
    # TV: the z-forward norm built from the y-backward diagonal twice
    "tv_zdiag_dup": (
        [("(xyzp - xmyzp) * (xyzp - xmyzp) +\n                                                  (xyzp - xymzp) * (xyzp - xymzp)",
          "(xyzp - xymzp) * (xyzp - xymzp) +\n                                                  (xyzp - xymzp) * (xyzp - xymzp)")],
        [(P.test_tv_gradient_is_gradient_of_tv_functional,)]),
    # general-3D BP: the two z weights exchanged between kL and kU
    "g3d_bp_z_weights": (
        [("R iZ = cur.z - bz - (R)kL + R(0.5);\n              R iZZ = R(1) - iZ;",
          "R iZZ = cur.z - bz - (R)kL + R(0.5);\n              R iZ = R(1) - iZZ;")],
        [(P.test_g3d_bp_quarter_voxel_z_profile,)]),
    # OSC-TV: TV from the first iteration on
    "osctv_tv_at_k1": (
        [("    if (k > 1) {\n      double sum = 0.0;", "    if (k >= 1) {\n      double sum = 0.0;")],
        [(P.test_osc_tv_recon_is_composition_of_parts, 2)]),
    # OSC-TV: element 1 of pre/post not zeroed for the later iterations
    "osctv_no_corner_zero": (
        [("  pre2[0] = (R)0;\n  post2[0] = (R)0;\n", "")],
        [(P.test_osc_tv_recon_is_composition_of_parts, 2)]),
    # OSC-TV: TV constant without the d_a factor
    "osctv_c_without_da": (
        [("(R)tv * d_a, tv_steps);", "(R)tv, tv_steps);")],
        [(P.test_osc_tv_recon_is_composition_of_parts, 1)]),
}


def _build(src_text, subs, out_dir):
    text = src_text
    for old, new in subs:
        n = text.count(old)
        assert n == 1, f"slip anchor found {n} times: {old!r}"
        text = text.replace(old, new)
    os.makedirs(out_dir, exist_ok=True)
    shutil.copy(os.path.join(ORACLE_DIR, "oracle.h"), out_dir)
    src = os.path.join(out_dir, "oracle.cpp")
    with open(src, "w") as f:
        f.write(text)
    lib = os.path.join(out_dir, "liboracle.so")
    subprocess.run(["g++", *CXXFLAGS, "-o", lib, src], check=True, capture_output=True)
    return lib


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = tmp_path_factory.mktemp("oracle_slips")
    with open(os.path.join(ORACLE_DIR, "oracle.cpp")) as f:
        text = f.read()
    jobs = {"unmodified": []}
    jobs.update({k: v[0] for k, v in SLIPS.items()})
    with cf.ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        futs = {k: ex.submit(_build, text, subs, str(root / k)) for k, subs in jobs.items()}
        return {k: f.result() for k, f in futs.items()}


def _pins(name):
    return SLIPS[name][1]


def test_unmodified_copy_passes_every_guarding_pin(builds):
    with O.use_library(builds["unmodified"]):
        for name in SLIPS:
            for fn, *args in _pins(name):
                fn(*args)


@pytest.mark.parametrize("name", sorted(SLIPS))
def test_pin_fails_on_planted_slip(builds, name):
    with O.use_library(builds[name]):
        for fn, *args in _pins(name):
            with pytest.raises(AssertionError):
                fn(*args)
