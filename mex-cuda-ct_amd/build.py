"""Build libmexct.so (the HIP kernels + C ABI) for gfx950 with hipcc.

    python mex-cuda-ct_amd/build.py            # compile what changed
    python mex-cuda-ct_amd/build.py --force    # recompile everything

Objects go to mex-cuda-ct_amd/_build/, the library to
mex-cuda-ct_amd/lib/libmexct.so.  Every translation unit is compiled with
-ffp-contract=off (one rounding per float operation, in source order) and
with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmexct.so")
SOURCES = ["pad.hip", "fp.hip", "fpz.hip", "bp.hip", "g3dbp.hip", "tv.hip", "fdk.hip", "probe.hip",
           "api.hip", "serve.hip"]
HEADERS = ["common.h", "internal.h", "ray.h"]
ARCH = os.environ.get("MEXCT_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def cflags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-munsafe-fp-atomics",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wno-unused-command-line-argument",
    ]


def _stale(obj: str, src: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "mexct.h")]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    if force or _stale(obj, path):
        cmd = [hipcc(), *cflags(), "-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l and l.strip())
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj


MEXDIR = os.path.join(HERE, "mex")
GATEWAYS = ["CUDAmex_FP", "CUDAmex_BP", "CUDAmex_general3D_FP", "CUDAmex_general3D_BP",
            "CUDAmex_oscIter", "CUDAmex_TVmin_3D"]


def build_mex(force: bool = False) -> list[str]:
    """Build the MEX gateways against the in-tree mex.h shim (MATLAB-free
    test build): lib/libmexshim.so and lib/mex/<name>.so, each exporting
    mexFunction and linked to libmexct.so."""
    gxx = shutil.which("g++") or "g++"
    shim_src = os.path.join(MEXDIR, "shim", "mex_shim.cpp")
    shim = os.path.join(LIBDIR, "libmexshim.so")
    inc = ["-I", os.path.join(MEXDIR, "shim"), "-I", MEXDIR, "-I", os.path.join(ROOT, "include")]
    if force or not os.path.exists(shim) or os.path.getmtime(shim_src) > os.path.getmtime(shim):
        subprocess.run([gxx, "-O2", "-std=c++17", "-fPIC", "-shared", *inc, shim_src, "-o", shim],
                       check=True)
    outdir = os.path.join(LIBDIR, "mex")
    os.makedirs(outdir, exist_ok=True)
    outs = []
    for g in GATEWAYS:
        src = os.path.join(MEXDIR, g + ".cpp")
        out = os.path.join(outdir, g + ".so")
        deps = [src, os.path.join(MEXDIR, "gateway.h"), os.path.join(ROOT, "include", "mexct.h")]
        if force or not os.path.exists(out) or any(os.path.getmtime(d) > os.path.getmtime(out) for d in deps):
            subprocess.run([gxx, "-O2", "-std=c++17", "-fPIC", "-shared", *inc, src, "-o", out,
                            "-L", LIBDIR, "-lmexct", "-lmexshim", "-Wl,-rpath,$ORIGIN/.."],
                           check=True)
        outs.append(out)
    return outs


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
        os.replace(tmp, LIB)
    build_mex(force)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.force, a.j))
    sys.exit(0)
