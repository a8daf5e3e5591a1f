"""Host build of the device exponential used for exp() in system sources
(sodecl_amd/csrc/device/exp.hip, compiled by g++ as C++ with contraction
off: the same IEEE double operations as on the GPU), and its error against
x87 long double expl (64-bit significand) measured in units in the last
place of the double result."""
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
#include "%s/exp.hip"
extern "C" void exp_host(const double* x, double* e, long n)
{
    for (long i = 0; i < n; ++i) e[i] = sodecl_exp(x[i]);
}
'''


@functools.lru_cache(None)
def _lib():
    d = tempfile.mkdtemp(prefix="sodecl_exp_")
    cpp = os.path.join(d, "exp_host.cpp")
    with open(cpp, "w") as f:
        f.write(_HARNESS % (DEVICE, DEVICE))
    so = os.path.join(d, "exp_host.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", cpp, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.exp_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return lib


def exp_host(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    e = np.empty_like(x)
    _lib().exp_host(x.ctypes.data, e.ctypes.data, x.size)
    return e


def ulp_error(x, e):
    """|e - exp x| in units of the spacing of doubles at exp x (normal results)."""
    ref = np.exp(np.asarray(x, dtype=np.longdouble))
    spacing = np.spacing(ref.astype(np.float64)).astype(np.longdouble)
    return np.asarray(np.abs(np.asarray(e, dtype=np.longdouble) - ref) / spacing, dtype=np.float64)
