"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

mex-cuda-ct_amd/sanitize.py builds the C ABI library (host side), the MEX
shim and gateways and the oracle with both sanitizers, and runs the CPU
tests of the C ABI's argument paths, the gateways and the oracle pins
against those builds in a child process.  Two planted mistakes, compiled
with the same flags and loaded the same way, show that a heap overflow and
a signed overflow end such a run with a report.
"""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mex-cuda-ct_amd")
HAVE = os.path.exists(os.path.join(PKG, "sanitize.py")) and os.path.exists("/opt/rocm/llvm/bin/clang++")
pytestmark = pytest.mark.skipif(not HAVE, reason="needs mex-cuda-ct_amd/sanitize.py and ROCm clang")


def _sanitize():
    if PKG not in sys.path:
        sys.path.insert(0, PKG)
    import sanitize
    return sanitize


@pytest.fixture(scope="module")
def san():
    S = _sanitize()
    S.build()
    return S


def test_cpu_tests_pass_under_asan_ubsan(san):
    r = san.run()
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in r.stdout


PLANTS = {
    # one element past a heap array
    "heap_overflow": ("extern \"C\" int plant(int n) { int* p = new int[n]; "
                      "for (int i = 0; i <= n; ++i) p[i] = i; int s = p[n / 2]; delete[] p; return s; }",
                      86, "heap-buffer-overflow"),
    # signed int overflow
    "signed_overflow": ("extern \"C\" int plant(int n) { int x = 2147483600; "
                        "for (int i = 0; i < n; ++i) x += 1000; return x; }",
                        86, "signed integer overflow"),
}


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_mistake_is_reported(san, name):
    src, code, text = PLANTS[name]
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "plant.cpp")
        so = os.path.join(d, "libplant.so")
        with open(cpp, "w") as f:
            f.write(src)
        subprocess.run([san.clangxx(), "-O0", "-g", "-fPIC", "-shared", *san.SAN_FLAGS, cpp, "-o", so],
                       check=True)
        prog = f"import ctypes; print(ctypes.CDLL({so!r}).plant(8))"
        r = subprocess.run([sys.executable, "-c", prog], env=san.env(), capture_output=True, text=True)
    assert r.returncode == code, (r.returncode, r.stderr[-2000:])
    assert text in r.stderr
