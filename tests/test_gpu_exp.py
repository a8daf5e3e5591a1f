"""exp() in a system source on the GPU: SODECL's exponential (the default)
against its host build bit for bit, and the device library's (library_exp =
1) within its own accuracy.

The values come out of the product path: a one-step Euler run with dt = 1
from y = -0 of the system f_i = exp(p_i) stores exp(p_i) exactly
(-0 + e = e for every e)."""
import numpy as np
import pytest

from sodecl_amd import sodecl_interface as si
from tests.exp_host import exp_host, ulp_error

pytestmark = pytest.mark.gpu

EXPS = """
void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
        yout[i] = exp(p[i]);
}
"""
N = 16


def device_exps(x, **options):
    assert x.size % N == 0
    orbits = x.size // N
    s = si.Session(EXPS, 1, orbits, N, N, 0, 1.0, 1.0, 1, 0, dict(schedule="lane", **options), True)
    s.setup(np.full((orbits, N), -0.0), x.reshape(orbits, N))
    out = s.run().reshape(2, N, orbits)
    np.testing.assert_array_equal(out[0], -0.0)
    return out[1].T.reshape(-1).copy()


def arguments():
    r = np.random.default_rng(12)
    parts = [r.uniform(-1, 1, 1 << 18), r.uniform(-708, 709.78, 1 << 18), r.uniform(-0.0014, 0.0014, 1 << 16),
             r.uniform(-745.2, -708, 1 << 14), np.arange(-2048, 2048) * (np.log(2) / 256),
             np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 709.78, 709.79, 1e300, -1e300, -745.13, -745.14,
                       5e-324, 1400.0, -1400.0, 1401.0, -1401.0])]
    x = np.concatenate(parts)
    pad = (-x.size) % N
    return np.concatenate([x, np.linspace(0.1, 1.0, pad)])


def test_exp_matches_host_build():
    x = arguments()
    got = device_exps(x)
    host = exp_host(x)
    nan = np.isnan(x)
    np.testing.assert_array_equal(got[~nan].view(np.uint64), host[~nan].view(np.uint64))
    assert np.isnan(got[nan]).all()
    normal = np.isfinite(x) & (x > -708) & (x < 709.78)
    assert ulp_error(x[normal], got[normal]).max() < 0.7


def test_library_exp_option():
    x = arguments()
    lib = device_exps(x, library_exp=1)
    own = device_exps(x)
    normal = np.isfinite(x) & (x > -708) & (x < 709.78)
    assert ulp_error(x[normal], lib[normal]).max() < 2.0
    assert not np.array_equal(lib[normal], own[normal])
    special = ~np.isnan(x) & ~normal
    np.testing.assert_array_equal(np.isfinite(lib[special]), np.isfinite(own[special]))
