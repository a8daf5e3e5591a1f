"""Host build of the device sine used for sin() in system sources
(sodecl_amd/csrc/device/sin.hip, compiled by g++ as C++ with contraction
off: the same IEEE double operations as on the GPU), and its error against
x87 long double sinl (64-bit significand) measured in units in the last
place of the double result."""
import ctypes
import functools
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "sodecl_amd", "csrc", "device", "sin.hip")

_HARNESS = r'''
#include <math.h>
#define SODECL_MATH_FN static inline
#include "%s"
extern "C" void sin_host(const double* x, double* s, unsigned char* library, long n)
{
    for (long i = 0; i < n; ++i) {
        bool l;
        s[i] = sodecl_sin_reduced(x[i], &l);
        library[i] = l;
    }
}
'''


@functools.lru_cache(None)
def _lib():
    d = tempfile.mkdtemp(prefix="sodecl_sin_")
    cpp = os.path.join(d, "sin_host.cpp")
    with open(cpp, "w") as f:
        f.write(_HARNESS % SRC)
    so = os.path.join(d, "sin_host.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", cpp, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.sin_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return lib


def sin_host(x):
    """(fast-path sine, flag: the device returns the library's sin instead)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    lib = np.empty(x.shape, dtype=np.uint8)
    _lib().sin_host(x.ctypes.data, s.ctypes.data, lib.ctypes.data, x.size)
    return s, lib.astype(bool)


def ulp_error(x, s):
    """|s - sin x| in units of the spacing of doubles at |sin x|."""
    ref = np.sin(np.asarray(x, dtype=np.longdouble))
    spacing = np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble)
    return np.asarray(np.abs(np.asarray(s, dtype=np.longdouble) - ref) / spacing, dtype=np.float64)
