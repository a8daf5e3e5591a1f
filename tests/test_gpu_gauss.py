"""The SDE noise on the GPU: Box-Muller pairs from the device's Threefry
blocks and its logarithm, sine and cosine (gauss.hip) against the host
build of the same functions bit for bit, and against the oracle's pairs
(host libm) to 2 ulp.

The values come out of the product path: one Euler-Maruyama step with
dt = 1 from y = 0 of the system f = 0, stoch_i = noise_i stores each noise
value exactly (0 + 0 dt + z = z)."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sodecl_amd import sodecl_interface as si
from tests.gauss_host import log_host, sincos_host

pytestmark = pytest.mark.gpu

NOISE = """
void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
        yout[i] = 0.0;
}
void sode_system_stoch(double t, double y[_numeq_], double stoch[_numeq_], double p[_numpar_],
                       double noise[_numnoi_])
{
    for (int i = 0; i < _numeq_; i++)
        stoch[i] = noise[i];
}
"""
N = 16           # noise values per orbit: 8 Threefry calls
ORBITS = 1024


def device_noise(counters):
    s = si.Session(NOISE, 0, ORBITS, N, 1, N, 1.0, 1.0, 1, 0, dict(schedule="lane"), True)
    s.setup(np.zeros((ORBITS, N)), np.zeros((ORBITS, 1)), None, counters)
    out = s.run().reshape(2, N, ORBITS)
    np.testing.assert_array_equal(s.counters(), counters + N // 2)
    return out[1].T   # [orbit][noise]


def host_noise(counters):
    kura = Oracle("kuramoto")
    u0 = np.empty((ORBITS, N // 2))
    u1 = np.empty((ORBITS, N // 2))
    libm = np.empty((ORBITS, N))
    for o in range(ORBITS):
        for k in range(N // 2):
            w0, w1 = kura.threefry((o, int(counters[o]) + k), (0, 0xF00DCAFE))
            u0[o, k], u1[o, k] = kura.u01(w0), kura.u01(w1)
            libm[o, 2 * k:2 * k + 2] = kura.noise_pair(o, int(counters[o]) + k, 1.0)
    r = np.sqrt(-2 * log_host(u0.ravel())).reshape(u0.shape)
    s, c = sincos_host((2 * np.pi * u1).ravel())
    own = np.empty((ORBITS, N))
    own[:, 0::2] = r * c.reshape(u1.shape)
    own[:, 1::2] = r * s.reshape(u1.shape)
    return own, libm


def test_noise_matches_host_build_and_oracle():
    counters = np.random.default_rng(21).integers(0, 2 ** 40, ORBITS).astype(np.float64)
    got = device_noise(counters)
    own, libm = host_noise(counters)
    np.testing.assert_array_equal(got.view(np.uint64), own.view(np.uint64))
    ulps = np.abs(got - libm) / np.spacing(np.abs(libm))
    assert ulps.max() <= 2.0
    assert np.mean(got == libm) > 0.97
