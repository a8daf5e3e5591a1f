"""ctypes access to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  Two back ends:

* ``Oracle(system)`` -- oracle/build/liboracle_<system>.so, the C restatement
  of the reference integrators (oracle/sodecl_oracle.c);
* ``RefKernels(system, solver)`` -- oracle/_ref/libref_<system>_<solver>.so,
  the reference's own OpenCL C kernels compiled as C (oracle/ref_driver.c),
  present only where the reference tree was available at build time.

Arrays follow the reference's host layouts: t0[orbits], y0[orbits][neq]
(AoS), params[orbits][npar], counters[orbits] (double), and the output
[nsnap][neq][orbits] of src/sodeclmgr.cpp:979-995.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
REF = os.path.join(HERE, "_ref")

SOLVERS = {"em": 0, "euler": 1, "rk4": 2, "ie": 3, "im": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


class _Config(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int),
        ("neq", ctypes.c_int),
        ("npar", ctypes.c_int),
        ("nnoi", ctypes.c_int),
        ("orbits", ctypes.c_int),
        ("ksteps", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("tspan", ctypes.c_double),
        ("newton_maxit", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("nlaunch", ctypes.c_long),
        ("use_band", ctypes.c_int),
        ("newton_tol", ctypes.c_double),
    ]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def num_launches(tspan, dt, ksteps):
    lib = ctypes.CDLL(os.path.join(BUILD, "liboracle_oculomotor.so"))
    lib.orc_num_launches.restype = ctypes.c_long
    lib.orc_num_launches.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
    return int(lib.orc_num_launches(tspan, dt, ksteps))


class Oracle:
    def __init__(self, system):
        path = os.path.join(BUILD, f"liboracle_{system}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_run.restype = ctypes.c_long
        L.orc_run.argtypes = [ctypes.POINTER(_Config), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _lp]
        L.orc_threefry2x64_20.argtypes = [ctypes.POINTER(ctypes.c_uint64)] * 3
        L.orc_u01_open_open_64_53.restype = ctypes.c_double
        L.orc_u01_open_open_64_53.argtypes = [ctypes.c_uint64]
        L.orc_kernel_dt.restype = ctypes.c_double
        L.orc_kernel_dt.argtypes = [ctypes.c_double]
        L.orc_num_launches.restype = ctypes.c_long
        L.orc_num_launches.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_noise_pair.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, _dp]
        L.orc_sensitivity_gradient.restype = ctypes.c_long
        L.orc_sensitivity_gradient.argtypes = [ctypes.POINTER(_Config), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                               _dp, _dp]
        L.orc_sample.argtypes = [ctypes.POINTER(ctypes.c_int), _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 _dp]
        L.orc_jacobians.restype = ctypes.c_long
        L.orc_jacobians.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_counter_word.restype = ctypes.c_uint64
        L.orc_counter_word.argtypes = [ctypes.c_double]

    def run(self, solver, y0, params, dt, tspan, ksteps, t0=None, counters=None, nnoi=0,
            newton_maxit=1000, orbit_range=None, nthreads=0, want_output=True, nlaunch=-1, use_band=False,
            newton_tol=0.0):
        """Integrates every orbit; returns dict(out, y, t, counters, newton_updates, launches).
        nlaunch >= 0 fixes the number of ksteps segments instead of tspan;
        newton_tol > 0 replaces the reference's Newton threshold 1e-07."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        params = np.ascontiguousarray(params, dtype=np.float64)
        orbits, neq = y0.shape
        npar = params.shape[1]
        c = _Config(SOLVERS[solver] if isinstance(solver, str) else int(solver), neq, npar, nnoi,
                    orbits, ksteps, dt, tspan, newton_maxit, nthreads, nlaunch, int(use_band), float(newton_tol))
        ob, oc = (0, orbits) if orbit_range is None else orbit_range
        t = np.zeros(orbits) if t0 is None else np.array(t0, dtype=np.float64)
        y = y0.copy()
        cnt = None if counters is None else np.array(counters, dtype=np.float64)
        launches = nlaunch if nlaunch >= 0 else self.lib.orc_num_launches(tspan, dt, ksteps)
        out = np.empty((launches + 1, neq, oc)) if want_output else None
        nu = np.zeros(orbits, dtype=np.int64)
        r = self.lib.orc_run(ctypes.byref(c), ob, oc, _ptr(t), _ptr(y), _ptr(params), _ptr(cnt),
                             _ptr(out), nu.ctypes.data_as(_lp))
        if r < 0:
            raise ValueError("oracle rejected the configuration")
        return dict(out=out, y=y, t=t, counters=cnt, newton_updates=nu, launches=int(r))

    def sensitivity_gradient(self, solver, y0, params, cot, dt, tspan, ksteps, orbit_range=None, nthreads=0,
                             newton_tol=0.0):
        """Gradient of sum <cot[snap], y(snap)> w.r.t. y0 and params by forward
        sensitivity; cot is [nsnap][neq][orbits].  Returns (grad_y0, grad_p)
        for the orbit range.  Implicit solvers: tangent of each step's
        implicit equation at its converged Newton solution."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        params = np.ascontiguousarray(params, dtype=np.float64)
        cot = np.ascontiguousarray(cot, dtype=np.float64)
        orbits, neq = y0.shape
        npar = params.shape[1]
        c = _Config(SOLVERS[solver], neq, npar, 0, orbits, ksteps, dt, tspan, 1000, nthreads, -1, 0, float(newton_tol))
        ob, oc = (0, orbits) if orbit_range is None else orbit_range
        gy = np.zeros((oc, neq))
        gp = np.zeros((oc, npar))
        r = self.lib.orc_sensitivity_gradient(ctypes.byref(c), ob, oc, _ptr(y0), _ptr(params), _ptr(cot), _ptr(gy),
                                              _ptr(gp))
        if r < 0:
            raise ValueError("sensitivity needs calc_jacobian/calc_jacobian_params and a deterministic solver")
        return gy, gp

    def jacobians(self, y, p, exact=True):
        """The plugin's (df/dy [neq][neq], df/dp [neq][npar]) at (y, p)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        jac = np.zeros((y.size, y.size))
        jp = np.zeros((y.size, max(p.size, 1)))
        if self.lib.orc_jacobians(y.size, p.size, _ptr(y), _ptr(p), _ptr(jac), _ptr(jp), int(exact)) < 0:
            raise ValueError("the plugin lacks calc_jacobian / calc_jacobian_params")
        return jac, jp[:, :p.size]

    def sample(self, table, rows, seed):
        """The on-device sampler restated: table is a list of (kind, a, b)."""
        kind = np.array([t[0] for t in table], dtype=np.int32)
        a = np.array([t[1] for t in table], dtype=np.float64)
        b = np.array([t[2] for t in table], dtype=np.float64)
        out = np.empty((rows, len(table)))
        self.lib.orc_sample(kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(a), _ptr(b), rows, len(table),
                            seed, _ptr(out))
        return out

    def threefry(self, ctr, key):
        c = (ctypes.c_uint64 * 2)(*ctr)
        k = (ctypes.c_uint64 * 2)(*key)
        o = (ctypes.c_uint64 * 2)()
        self.lib.orc_threefry2x64_20(c, k, o)
        return int(o[0]), int(o[1])

    def u01(self, x):
        return self.lib.orc_u01_open_open_64_53(x)

    def kernel_dt(self, dt):
        return self.lib.orc_kernel_dt(dt)

    def rhs(self, y, p, t=0.0, nnoi=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty_like(y)
        self.lib.orc_rhs(y.size, p.size, nnoi, t, _ptr(y), _ptr(out), _ptr(p))
        return out

    def noise_pair(self, orbit, counter_word, dt):
        out = np.empty(2)
        self.lib.orc_noise_pair(orbit, counter_word, dt, _ptr(out))
        return out


class RefKernels:
    """The reference's kernels compiled as C; run() mirrors run_sode_solver.
    fast: the -O3 scalar build; simd: the orbit-vectorised build
    (oracle/ref_simd.c, compile-time sizes, deterministic solvers only)."""

    def __init__(self, system, solver, fast=False, simd=False):
        suffix = "_simd" if simd else ("_fast" if fast else "")
        path = os.path.join(REF, f"libref_{system}_{solver}{suffix}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.lib.ref_run.restype = ctypes.c_long
        self.lib.ref_run.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 2 + [_dp] * 5 + [ctypes.c_int]
        self.stochastic = solver == "em"

    def run(self, y0, params, dt, tspan, ksteps, t0=None, counters=None, nnoi=0, nthreads=0,
            want_output=True):
        y = np.array(y0, dtype=np.float64, order="C")
        p = np.ascontiguousarray(params, dtype=np.float64)
        orbits, neq = y.shape
        t = np.zeros(orbits) if t0 is None else np.array(t0, dtype=np.float64)
        cnt = None if counters is None else np.array(counters, dtype=np.float64)
        if self.stochastic and cnt is None:
            raise ValueError("stochastic run needs counters")
        launches = num_launches(tspan, dt, ksteps)
        out = np.empty((launches + 1, neq, orbits)) if want_output else None
        r = self.lib.ref_run(neq, p.shape[1], nnoi, orbits, ksteps, dt, tspan, _ptr(t), _ptr(y),
                             _ptr(p), _ptr(cnt), _ptr(out), nthreads)
        if r < 0:
            raise ValueError("bad configuration")
        return dict(out=out, y=y, t=t, counters=cnt, launches=int(r))


def have_ref(system="oculomotor", solver="rk4", simd=False):
    return os.path.exists(os.path.join(REF, f"libref_{system}_{solver}{'_simd' if simd else ''}.so"))
