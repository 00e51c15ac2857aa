"""Call the MEX gateways (mex-cuda-ct_amd/lib/mex/*.so, built against the
in-tree mex.h shim) from Python the way MATLAB would: numpy arrays become
mxArrays, mexFunction runs, errors come back as MexCallError(id, msg)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MEXCT_LIB_DIR: another build of the shim and gateways (the sanitizer build)
LIBDIR = os.environ.get("MEXCT_LIB_DIR") or os.path.join(ROOT, "mex-cuda-ct_amd", "lib")

CLASS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.int32): 12,
         np.dtype(np.uint8): 9}
DTYPE = {v: k for k, v in CLASS.items()}

_shim = None
_gw = {}


class MexCallError(RuntimeError):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident
        self.msg = msg


def shim():
    global _shim
    if _shim is None:
        _shim = C.CDLL(os.path.join(LIBDIR, "libmexshim.so"), mode=C.RTLD_GLOBAL)
        _shim.shim_new.restype = C.c_void_p
        _shim.shim_new.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_size_t), C.c_void_p]
        _shim.shim_free.argtypes = [C.c_void_p]
        _shim.shim_invoke.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        _shim.shim_last_id.restype = C.c_char_p
        _shim.shim_last_msg.restype = C.c_char_p
        for f in ("mxGetData",):
            getattr(_shim, f).restype = C.c_void_p
            getattr(_shim, f).argtypes = [C.c_void_p]
        _shim.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        _shim.mxGetDimensions.argtypes = [C.c_void_p]
        _shim.mxGetNumberOfDimensions.restype = C.c_size_t
        _shim.mxGetNumberOfDimensions.argtypes = [C.c_void_p]
        _shim.mxGetClassID.restype = C.c_int
        _shim.mxGetClassID.argtypes = [C.c_void_p]
    return _shim


def gateway(name):
    if name not in _gw:
        shim()
        lib = C.CDLL(os.path.join(LIBDIR, "mex", name + ".so"))
        _gw[name] = (lib, C.cast(lib.mexFunction, C.c_void_p))
    return _gw[name][1]


def to_mx(a):
    a = np.asarray(a)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    a = np.asfortranarray(a)
    dims = (C.c_size_t * a.ndim)(*a.shape)
    return shim().shim_new(CLASS[a.dtype], a.ndim, dims, a.ctypes.data_as(C.c_void_p))


def from_mx(p):
    s = shim()
    nd = s.mxGetNumberOfDimensions(p)
    dims = tuple(s.mxGetDimensions(p)[i] for i in range(nd))
    dt = DTYPE[s.mxGetClassID(p)]
    n = int(np.prod(dims))
    buf = (C.c_char * (n * dt.itemsize)).from_address(s.mxGetData(p))
    return np.frombuffer(buf, dtype=dt).reshape(dims, order="F").copy()


def call(name, *args, nlhs=1, inplace=()):
    """Run mexFunction of gateway `name`; returns the outputs (a list) and,
    for the indices in `inplace`, copies the mutated inputs back."""
    s = shim()
    fn = gateway(name)
    prhs = [to_mx(a) for a in args]
    parr = (C.c_void_p * max(len(prhs), 1))(*prhs)
    plhs = (C.c_void_p * max(nlhs, 1))()
    try:
        rc = s.shim_invoke(fn, nlhs, plhs, len(prhs), parr)
        if rc != 0:
            raise MexCallError(s.shim_last_id().decode(), s.shim_last_msg().decode())
        outs = [from_mx(plhs[i]) for i in range(nlhs) if plhs[i]]
        for i in inplace:
            args[i][...] = from_mx(prhs[i])
        return outs
    finally:
        for p in prhs:
            s.shim_free(p)
        for i in range(nlhs):
            if plhs[i]:
                s.shim_free(plhs[i])
