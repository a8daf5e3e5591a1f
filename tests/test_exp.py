"""The exponential that system sources get for exp() of a double
(sodecl_amd/csrc/device/exp.hip), checked on the CPU.

The device function is plain IEEE double arithmetic (fma, products, sums,
ldexp and a table), so its host build computes the same bits as the GPU
(tests/test_gpu_exp.py compares them); here that build is measured against
x87 long double expl (64-bit significand, error far below a double ulp),
against exp(a + b) = exp(a) exp(b) for exact power-of-two steps, the
overflow and underflow thresholds, the subnormal range and the special
values.  The derivation script's table and constants are checked against
the source."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from tests.exp_host import DEVICE, exp_host, ulp_error

RANGES = [
    ("|x| <= ln2/512", lambda r, n: r.uniform(-0.00135, 0.00135, n)),
    ("|x| <= 1", lambda r, n: r.uniform(-1, 1, n)),
    ("|x| <= 30", lambda r, n: r.uniform(-30, 30, n)),
    ("-708..0", lambda r, n: r.uniform(-708, 0, n)),
    ("0..709", lambda r, n: r.uniform(0, 709.78, n)),
    ("log-uniform 1e-300..700, both signs", lambda r, n: np.exp(r.uniform(np.log(1e-300), np.log(700), n))
     * r.choice([-1.0, 1.0], n)),
]


def _long_double_ok():
    return np.finfo(np.longdouble).nmant >= 63


@pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double for the reference exponential")
@pytest.mark.parametrize("name,draw", RANGES, ids=[r[0] for r in RANGES])
def test_exp_error_below_0_7_ulp(name, draw):
    x = draw(np.random.default_rng(3), 400_000)
    e = exp_host(x)
    err = ulp_error(x, e)
    assert err.max() < 0.7, f"{name}: {err.max():.3f} ulp at {x[np.argmax(err)]!r}"
    assert np.mean(err > 0.5) < 0.02, name


def test_exp_special_values_and_thresholds():
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 709.78, 709.79, 710.0, 1e300, -745.13, -745.14, -746.0,
                  -1e300, 5e-324, -5e-324, 1e-20, 1400.0, -1400.0, 1401.0, -1401.0])
    e = exp_host(x)
    assert e[0] == 1.0 and e[1] == 1.0
    assert e[2] == np.inf and e[3] == 0.0 and np.isnan(e[4])
    assert np.isfinite(e[5]) and e[6] == np.inf and e[7] == np.inf and e[8] == np.inf
    assert e[9] == 5e-324 and e[10] == 0.0 and e[11] == 0.0 and e[12] == 0.0
    assert e[13] == 1.0 and e[14] == 1.0 and e[15] == 1.0
    assert e[16] == np.inf and e[17] == 0.0 and e[18] == np.inf and e[19] == 0.0
    # NaN payloads and signs stay NaN
    assert np.isnan(exp_host(np.array([-np.nan, np.float64("nan")]))).all()


@pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double")
def test_exp_subnormal_results():
    """Gradual underflow through ldexp: results in the subnormal range are
    within one unit of the exact value rounded to that range."""
    x = np.random.default_rng(4).uniform(-745.1, -708.4, 200_000)
    e = exp_host(x)
    ref = np.exp(x.astype(np.longdouble))
    assert np.all(np.abs(e.astype(np.longdouble) - ref) <= np.longdouble(5e-324))


@pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double")
def test_exp_power_of_two_steps():
    """exp(x + k ln2) = 2^k exp(x): with x on the table grid j ln2/256 (plus a
    small offset) for every j, the table points and the scaling by 2^m agree
    with each other to rounding (the exact ratio accounts for x + k ln2 being
    rounded to double)."""
    j = np.arange(256)
    base = j * (np.log(2) / 256) + 1e-5
    e0 = exp_host(base).astype(np.longdouble)
    ln2 = np.log(np.longdouble(2))
    for k in (-900, -40, -1, 1, 7, 300):
        xk = base + k * np.log(2)
        ek = exp_host(xk).astype(np.longdouble)
        want = np.exp((xk.astype(np.longdouble) - base.astype(np.longdouble)) - k * ln2)
        got = ek / (e0 * np.ldexp(np.longdouble(1), k))
        assert np.all(np.abs(got / want - 1) < 3e-16), k


def test_exp_constants_match_derivation():
    """The table and constants in exp.hip are those exp_coefficients.py
    derives (mpmath at 60 digits)."""
    pytest.importorskip("mpmath")
    out = subprocess.run([sys.executable, os.path.join(DEVICE, "exp_coefficients.py"), "--table"],
                         check=True, capture_output=True, text=True).stdout
    src = open(os.path.join(DEVICE, "exp.hip")).read()
    table_lines = [l.strip() for l in out.splitlines() if l.startswith("    ")]
    assert len(table_lines) == 128
    for line in table_lines:
        assert line in src
    for name in ("256/ln2", "C1", "C2"):
        value = re.search(rf"^{re.escape(name)}\s+(\S+)", out, re.M).group(1)
        assert value in src, name
    for q in re.search(r"^Q\s+\[(.*)\]", out, re.M).group(1).replace("'", "").split(", "):
        if q != "0x1.0000000000000p-1":
            assert q in src, q


@pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double")
def test_dropping_the_table_remainder_fails_the_bound(tmp_path):
    """The 0.7 ulp bound is tight enough to notice the table's second word:
    the same source with T_lo dropped (a copy built here, with
    fma(T_hi, p, 0) in place of fma(T_hi, p, T_lo)) exceeds it."""
    import ctypes
    src = open(os.path.join(DEVICE, "exp.hip")).read()
    assert src.count("fma(thi, p, tlo)") == 1
    (tmp_path / "exp.hip").write_text(src.replace("fma(thi, p, tlo)", "fma(thi, p, 0.0)"))
    cpp = tmp_path / "h.cpp"
    cpp.write_text('#include <math.h>\n#define SODECL_MATH_FN static inline\n'
                   f'#include "{DEVICE}/sin.hip"\n#include "{tmp_path}/exp.hip"\n'
                   'extern "C" void f(const double* x, double* e, long n)'
                   '{ for (long i = 0; i < n; ++i) e[i] = sodecl_exp(x[i]); }\n')
    so = tmp_path / "h.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", str(cpp), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    x = np.random.default_rng(9).uniform(-700, 700, 400_000)
    e = np.empty_like(x)
    lib.f(x.ctypes.data, e.ctypes.data, x.size)
    assert ulp_error(x, e).max() > 0.7
