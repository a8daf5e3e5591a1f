"""sin() in a system source on the GPU: SODECL's sine (the default) against
its host build bit for bit, and against the device library's sine
(library_sin = 1) where it hands over to it.

The values come out of the product path: a one-step Euler run with dt = 1
from y = -0 of the system f_i = sin(p_i) stores sin(p_i) exactly
(-0 + s = s for every s, zeros included)."""
import numpy as np
import pytest

from sodecl_amd import sodecl_interface as si
from tests.sin_host import sin_host, ulp_error

pytestmark = pytest.mark.gpu

SINES = """
void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
        yout[i] = sin(p[i]);
}
"""
N = 16


def device_sines(x, **options):
    assert x.size % N == 0
    orbits = x.size // N
    s = si.Session(SINES, 1, orbits, N, N, 0, 1.0, 1.0, 1, 0, dict(schedule="lane", **options), True)
    s.setup(np.full((orbits, N), -0.0), x.reshape(orbits, N))
    out = s.run().reshape(2, N, orbits)
    np.testing.assert_array_equal(out[0], -0.0)
    return out[1].T.reshape(-1).copy()


def arguments():
    r = np.random.default_rng(11)
    parts = [r.uniform(-np.pi / 2, np.pi / 2, 1 << 18), np.pi / 2 + r.uniform(-0.3, 0.3, 1 << 16),
             r.uniform(-100, 100, 1 << 18), r.uniform(-2.0 ** 30, 2.0 ** 30, 1 << 16),
             np.exp(r.uniform(np.log(1e-300), np.log(1e9), 1 << 16)) * r.choice([-1.0, 1.0], 1 << 16),
             np.arange(1, 4097) * np.pi, -np.arange(1, 4097) * np.pi,
             np.array([0.0, -0.0, 5e-324, -5e-324, 1e-300, np.pi / 2, -np.pi / 2, 2.0 ** 30, -(2.0 ** 30),
                       2.0 ** 31, -1e300, np.inf, -np.inf, np.nan, 1e22, 3.0])]
    x = np.concatenate(parts)
    pad = (-x.size) % N
    return np.concatenate([x, np.linspace(0.1, 1.0, pad)])


def test_sin_matches_host_build_and_library():
    x = arguments()
    got = device_sines(x)
    lib = device_sines(x, library_sin=1)
    host, fallback = sin_host(x)
    fast = ~fallback
    assert fallback.sum() >= 8192 + 6
    # fast path: the host build's bits; elsewhere the device library's
    np.testing.assert_array_equal(got[fast].view(np.uint64), host[fast].view(np.uint64))
    np.testing.assert_array_equal(got[fallback].view(np.uint64), lib[fallback].view(np.uint64))
    finite = np.isfinite(x)
    assert ulp_error(x[fast & finite], got[fast & finite]).max() < 0.9
    assert ulp_error(x[finite], lib[finite]).max() < 2.0
    assert np.isnan(got[~finite]).all()
    zero = np.flatnonzero((x == 0) & np.signbit(x))
    assert np.signbit(got[zero]).all() and (got[zero] == 0).all()
