"""Host build of the Box-Muller transform's logarithm, sine and cosine
(sodecl_amd/csrc/device/gauss.hip, compiled by g++ as C++ with contraction
off: the same IEEE double operations as on the GPU), and their errors
against x87 long double (64-bit significand) in units in the last place of
the double result."""
import ctypes
import functools
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEVICE = os.path.join(ROOT, "sodecl_amd", "csrc", "device")

_HARNESS = r'''
#include <math.h>
#define SODECL_MATH_FN static inline
#include "%s/sin.hip"
#include "%s/gauss.hip"
extern "C" void log_host(const double* x, double* y, long n)
{
    for (long i = 0; i < n; ++i) y[i] = sodecl_log_positive(x[i]);
}
extern "C" void sincos_host(const double* x, double* s, double* c, long n)
{
    for (long i = 0; i < n; ++i) sodecl_sincos_reduced(x[i], s + i, c + i);
}
'''


def build(device_dir=DEVICE, out_dir=None):
    d = out_dir or tempfile.mkdtemp(prefix="sodecl_gauss_")
    cpp = os.path.join(d, "gauss_host.cpp")
    with open(cpp, "w") as f:
        f.write(_HARNESS % (DEVICE, device_dir))
    so = os.path.join(d, "gauss_host.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", cpp, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.log_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.sincos_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return lib


@functools.lru_cache(None)
def _lib():
    return build()


def log_host(x, lib=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    (lib or _lib()).log_host(x.ctypes.data, y.ctypes.data, x.size)
    return y


def sincos_host(x, lib=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    (lib or _lib()).sincos_host(x.ctypes.data, s.ctypes.data, c.ctypes.data, x.size)
    return s, c


def ulp_error(got, ref):
    """|got - ref| in units of the spacing of doubles at ref (ref in long double)."""
    spacing = np.spacing(np.abs(ref).astype(np.float64)).astype(np.longdouble)
    return np.asarray(np.abs(np.asarray(got, dtype=np.longdouble) - ref) / spacing, dtype=np.float64)


def correctly_rounded(ref):
    """ref (long double) rounded to the nearest double."""
    return np.asarray(ref, dtype=np.float64)
