"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code, and
a run of the CPU tests against them.

    python mex-cuda-ct_amd/sanitize.py build   # builds into _san/ dirs
    python mex-cuda-ct_amd/sanitize.py run     # CPU tests on the _san builds

What is instrumented (host code only; no GPU sanitizer is used):
- mex-cuda-ct_amd/_san/libmexct.so: every host function of the C ABI
  (argument checks, workspace and table set-up, the request server), the
  device code compiled as usual;
- mex-cuda-ct_amd/_san/libmexshim.so and _san/mex/<gateway>.so: the MEX
  shim and the six gateways;
- oracle/_san/liboracle.so: the CPU oracle.

Everything is compiled by ROCm's clang with -shared-libsan, so one clang
ASan runtime serves the whole process.  `run` starts pytest in a child
process with that runtime added in front of LD_PRELOAD and the library
paths pointed at the _san builds (MEXCT_LIB, MEXCT_LIB_DIR,
MEXCT_ORACLE_LIB); any sanitizer report fails the run
(halt_on_error, -fno-sanitize-recover).  Leak checks are off: the HIP
runtime and Python keep allocations until exit by design.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

import build as _b

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "_san")
ORACLE = os.path.join(ROOT, "oracle")
ORACLE_SAN = os.path.join(ORACLE, "_san")
LLVM = "/opt/rocm/llvm/bin"

SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-shared-libsan"]
# the CPU tests that exercise the host code: the C ABI's argument paths, the
# MEX gateways through the shim, and the oracle with its pins
TESTS = ["tests/test_cpu_abi.py", "tests/test_mex_gateways.py", "tests/test_oracle_pins.py"]


def clangxx() -> str:
    return os.path.join(LLVM, "clang++")


def runtime() -> str:
    """The shared clang ASan runtime (UBSan included)."""
    r = subprocess.run([clangxx(), "-print-resource-dir"], capture_output=True, text=True,
                       check=True).stdout.strip()
    p = os.path.join(r, "lib", "linux", "libclang_rt.asan-x86_64.so")
    if not os.path.exists(p):
        raise RuntimeError(f"no ASan runtime at {p}")
    return p


def _host_only(flags):
    """Each sanitizer flag for the host pass of hipcc only."""
    out = []
    for f in flags:
        out += ["-Xarch_host", f]
    return out


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:3])} ... failed:\n{r.stdout}{r.stderr}")


def build() -> None:
    os.makedirs(os.path.join(SAN, "obj"), exist_ok=True)
    os.makedirs(os.path.join(SAN, "mex"), exist_ok=True)
    os.makedirs(ORACLE_SAN, exist_ok=True)
    host = _host_only(SAN_FLAGS[:-1])  # -shared-libsan is a link flag

    def compile_one(src):
        obj = os.path.join(SAN, "obj", src.replace(".hip", ".o"))
        _run([_b.hipcc(), *_b.cflags(), "-O1", *host, "-c", os.path.join(_b.CSRC, src), "-o", obj])
        return obj
    with cf.ThreadPoolExecutor(min(len(_b.SOURCES), os.cpu_count() or 2)) as ex:
        objs = list(ex.map(compile_one, _b.SOURCES))
    lib = os.path.join(SAN, "libmexct.so")
    _run([_b.hipcc(), f"--offload-arch={_b.ARCH}", "-shared", "-fPIC",
          *_host_only(SAN_FLAGS), "-o", lib, *objs])
    inc = ["-I", os.path.join(_b.MEXDIR, "shim"), "-I", _b.MEXDIR, "-I", os.path.join(ROOT, "include")]
    cx = [clangxx(), "-O1", "-g", "-std=c++17", "-fPIC", "-shared", *SAN_FLAGS]
    shim = os.path.join(SAN, "libmexshim.so")
    _run([*cx, *inc, os.path.join(_b.MEXDIR, "shim", "mex_shim.cpp"), "-o", shim])
    for g in _b.GATEWAYS:
        _run([*cx, *inc, os.path.join(_b.MEXDIR, g + ".cpp"), "-o",
              os.path.join(SAN, "mex", g + ".so"), "-L", SAN, "-lmexct", "-lmexshim",
              "-Wl,-rpath," + SAN])
    _run([*cx, "-pthread", "-ffp-contract=off", "-fno-fast-math", os.path.join(ORACLE, "oracle.cpp"),
          "-o", os.path.join(ORACLE_SAN, "liboracle.so")])


def env() -> dict:
    e = dict(os.environ)
    pre = e.get("LD_PRELOAD", "")
    e["LD_PRELOAD"] = runtime() + (" " + pre if pre else "")
    e["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86"
    e["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=86"
    e["MEXCT_LIB"] = os.path.join(SAN, "libmexct.so")
    e["MEXCT_LIB_DIR"] = SAN
    e["MEXCT_ORACLE_LIB"] = os.path.join(ORACLE_SAN, "liboracle.so")
    return e


def run(extra=()) -> subprocess.CompletedProcess:
    """pytest over TESTS (CPU only) on the sanitized builds, in a child."""
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
           *TESTS, *extra]
    return subprocess.run(cmd, cwd=ROOT, env=env(), capture_output=True, text=True)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "build"
    if what == "build":
        build()
        print(SAN)
    else:
        r = run(sys.argv[2:])
        print(r.stdout[-4000:], r.stderr[-4000:])
        sys.exit(r.returncode)
