"""ctypes wrapper of oracle/lib/libcpu_ref.so (test infrastructure only)."""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcpu_ref.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_cpu
            build_cpu.build(verbose=False)
        L = ctypes.CDLL(LIB_PATH)
        P, L_, I, D = ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_double
        L.cpu_applyK.argtypes = [P, L_, P, L_, I, P, I, P, I, D, P]
        L.cpu_applyK.restype = None
        L.cpu_applyDiffKT.argtypes = [P, L_, P, L_, I, P, P, I, I, D, P, I, D, P]
        L.cpu_applyDiffKT.restype = None
        L.cpu_threads.restype = I
        _lib = L
    return _lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def threads():
    return lib().cpu_threads()


def applyK(y, x, a, scale, KP=-1):
    y, x, a = _c(y), _c(x), _c(a)
    a2 = a.reshape(x.shape[0], -1)
    s = _c(np.atleast_1d(scale))
    out = np.zeros((y.shape[0], a2.shape[1]))
    lib().cpu_applyK(y.ctypes.data, y.shape[0], x.ctypes.data, x.shape[0], y.shape[1],
                     a2.ctypes.data, a2.shape[1], s.ctypes.data, s.size, float(KP), out.ctypes.data)
    return out


def applyDiffKT(y, x, p, a, scale, regweight=1., sym=False, lddmm=False, KP=-1):
    y, x, p, a = _c(y), _c(x), _c(p), _c(a)
    s = _c(np.atleast_1d(scale))
    mode = 2 if lddmm else (1 if sym else 0)
    out = np.zeros(y.shape)
    lib().cpu_applyDiffKT(y.ctypes.data, y.shape[0], x.ctypes.data, x.shape[0], y.shape[1],
                          p.ctypes.data, a.ctypes.data, a.shape[1], mode, float(regweight),
                          s.ctypes.data, s.size, float(KP), out.ctypes.data)
    return out
