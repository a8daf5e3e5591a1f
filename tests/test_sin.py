"""The sine that system sources get for sin() of a double
(sodecl_amd/csrc/device/sin.hip), checked on the CPU.

The device function is plain IEEE double arithmetic (rint, fma, products and
sums, no contraction), so its host build computes the same bits as the GPU
(tests/test_gpu_sin.py compares them); here that build is measured against
x87 long double sinl (64-bit significand, error far below a double ulp) and
against the identities sin(-x) = -sin(x), sin(x + k pi) = (-1)^k sin(x)
near exact multiples, and sin(x) = x for tiny x."""
import numpy as np
import pytest

import sodecl_amd
from sodecl_amd import sodecl_interface as si
from tests.sin_host import sin_host, ulp_error

PI = np.pi
RANGES = [
    ("|x| <= pi/2", lambda r, n: r.uniform(-PI / 2, PI / 2, n)),
    ("around pi/2", lambda r, n: PI / 2 + r.uniform(-0.3, 0.3, n)),
    ("|x| <= 10", lambda r, n: r.uniform(-10, 10, n)),
    ("|x| <= 1e3", lambda r, n: r.uniform(-1e3, 1e3, n)),
    ("|x| <= 2^30", lambda r, n: r.uniform(-2.0 ** 30, 2.0 ** 30, n)),
    ("log-uniform 1e-300..1e9", lambda r, n: np.exp(r.uniform(np.log(1e-300), np.log(1e9), n))
     * r.choice([-1.0, 1.0], n)),
]


def _long_double_ok():
    return np.finfo(np.longdouble).nmant >= 63


@pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double for the reference sine")
@pytest.mark.parametrize("name,draw", RANGES, ids=[r[0] for r in RANGES])
def test_sin_error_below_one_ulp(name, draw):
    x = draw(np.random.default_rng(7), 400_000)
    s, library = sin_host(x)
    fast = ~library
    assert library.mean() < 1e-4, name
    err = ulp_error(x[fast], s[fast])
    assert err.max() < 0.9, f"{name}: {err.max():.3f} ulp"
    # mostly correctly rounded
    assert np.mean(err > 0.5) < 0.1, name


def test_sin_odd_and_exact_cases():
    x = np.random.default_rng(8).uniform(0, 1e4, 100_000)
    s, lib = sin_host(x)
    sn, libn = sin_host(-x)
    np.testing.assert_array_equal(lib, libn)
    np.testing.assert_array_equal(sn[~lib], -s[~lib])
    tiny = np.array([1e-300, -1e-300, 5e-324, -5e-324, 1e-20, -3e-9, 2.0 ** -27])
    st, libt = sin_host(tiny)
    assert not libt.any()
    np.testing.assert_array_equal(st, tiny)
    z, libz = sin_host(np.array([0.0, -0.0]))
    assert not libz.any()
    assert z[0] == 0 and not np.signbit(z[0])
    assert z[1] == 0 and np.signbit(z[1])
    one, _ = sin_host(np.array([PI / 2, -PI / 2]))
    np.testing.assert_array_equal(one, [1.0, -1.0])


def test_sin_falls_back_where_the_reduction_is_short():
    """The library's sine is used for NaN, infinities, |x| > 2^30 and
    arguments within 4 |k P2| of k pi, and only there."""
    specials = np.array([np.nan, np.inf, -np.inf, 2.0 ** 30 * 1.0000001, -1e300])
    _, lib = sin_host(specials)
    assert lib.all()
    k = np.arange(1, 200_001, dtype=np.float64)
    multiples = k * PI                        # each within half an ulp of k pi
    _, lib = sin_host(np.concatenate([multiples, -multiples]))
    assert lib.all()
    off = multiples + 1e-9 * np.maximum(1.0, k * 1e-4)
    s, lib = sin_host(off)
    assert not lib.any()
    assert ulp_error(off, s).max() < 0.9
    _, lib = sin_host(np.array([2.0 ** 30, -(2.0 ** 30)]))
    assert not lib.any()


@pytest.mark.parametrize("library_sin", [0, 1])
def test_program_maps_sin_in_the_system_source_only(library_sin):
    s = si.Session(sodecl_amd.system_path("kuramoto"), 2, 64, 10, 11, 0, 0.01, 1.0, 10, 0,
                   {"schedule": "group", "library_sin": library_sin})
    s.compile_only("gfx950")
    src = s.program_source()
    if library_sin:
        assert "#define sin sodecl_sin" not in src
    else:
        system = src.index('#line 1 "system.cl"')
        assert src.index("#define sin sodecl_sin") < system < src.index("#undef sin") < \
            src.index('#line 1 "integrators.hip"')


def test_library_sin_option_is_validated():
    s = si.Session(sodecl_amd.system_path("kuramoto"), 2, 64, 10, 11, 0, 0.01, 1.0, 10, 0, {})
    with pytest.raises(RuntimeError, match="library_sin"):
        s.set_options({"library_sin": 2})


def test_sin_overloads_compile():
    """sin() of a float keeps single precision, sin() of an integer converts
    to double, as with the library's overloads."""
    src = """
void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    float a = sin(0.5f);
    int n = 2;
    yout[0] = sin(y[0]) + (double)a + sin(n) + sin(1L);
}
"""
    s = si.Session(src, 1, 64, 1, 1, 0, 0.1, 1.0, 1, 0, {}, True)
    assert s.compile_only("gfx950") > 1000
