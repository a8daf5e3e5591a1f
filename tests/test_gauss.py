"""The Box-Muller transform's logarithm, sine and cosine
(sodecl_amd/csrc/device/gauss.hip), checked on the CPU.

The device functions are plain IEEE double arithmetic with contraction off,
so their host build computes the same bits as the GPU (tests/test_gpu_gauss.py
compares them through the product path).  Here that build is measured
against x87 long double logl/sinl/cosl (64-bit significand, far below a
double ulp) on the arguments the uniforms produce, on the reduction's edges
(u near 1 and near the sqrt(1/2) split, x near multiples of pi/2), and
against identities that hold exactly; copies of the source with one term
dropped must fail the bounds, and the constants must be the derivation
script's."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from tests.gauss_host import DEVICE, build, correctly_rounded, log_host, sincos_host, ulp_error

U01_SCALE = float.fromhex("0x1.fffffffffffffp-54")


def _long_double_ok():
    return np.finfo(np.longdouble).nmant >= 63


pytestmark = pytest.mark.skipif(not _long_double_ok(), reason="needs x87 long double for the references")


def uniforms(n, seed):
    """The open-open uniforms of 64-bit words, as the noise draws them."""
    r = np.random.default_rng(seed)
    w = r.integers(0, 2 ** 63, size=n, dtype=np.uint64) * np.uint64(2) + r.integers(0, 2, size=n, dtype=np.uint64)
    return (1.0 + (w >> np.uint64(11)).astype(np.float64)) * U01_SCALE


def log_edges():
    one = 1.0 - np.arange(1, 2001) * 2.0 ** -53                  # just below 1
    split = float.fromhex("0x1.6a09e667f3bcdp-1") * (1 + np.arange(-1000, 1001) * 2.0 ** -52)   # around sqrt(1/2)
    powers = 2.0 ** -np.arange(1, 54)
    low = (1.0 + np.arange(0, 2000)) * U01_SCALE                   # the smallest uniforms
    return np.concatenate([one, split, split / 2 ** 7, powers, low])


def test_log_error_on_the_uniforms():
    u = uniforms(1_000_000, 1)
    y = log_host(u)
    ref = np.log(u.astype(np.longdouble))
    err = ulp_error(y, ref)
    assert err.max() < 0.66, f"{err.max():.3f} ulp at {u[np.argmax(err)]!r}"
    assert np.mean(y == correctly_rounded(ref)) > 0.99


def test_log_edges_and_exact_values():
    u = log_edges()
    y = log_host(u)
    err = ulp_error(y, np.log(u.astype(np.longdouble)))
    assert err.max() < 0.66, f"{err.max():.3f} ulp at {u[np.argmax(err)]!r}"
    # log 2^-k = -k ln2 rounded once; log 1 = 0
    k = np.arange(1, 54)
    want = correctly_rounded(-k * np.log(np.longdouble(2)))
    np.testing.assert_array_equal(log_host(2.0 ** -k), want)
    assert log_host(np.array([1.0]))[0] == 0.0
    # the whole positive normal range, not only the uniforms'
    x = np.exp(np.random.default_rng(2).uniform(np.log(1e-300), np.log(1e300), 200_000))
    assert ulp_error(log_host(x), np.log(x.astype(np.longdouble))).max() < 0.66


def test_sincos_error_on_the_angles():
    x = 2 * np.pi * uniforms(1_000_000, 3)
    s, c = sincos_host(x)
    xl = x.astype(np.longdouble)
    rs, rc = np.sin(xl), np.cos(xl)
    es, ec = ulp_error(s, rs), ulp_error(c, rc)
    assert es.max() < 0.6, f"sin {es.max():.3f} ulp at {x[np.argmax(es)]!r}"
    assert ec.max() < 0.6, f"cos {ec.max():.3f} ulp at {x[np.argmax(ec)]!r}"
    assert np.mean(s == correctly_rounded(rs)) > 0.99
    assert np.mean(c == correctly_rounded(rc)) > 0.99


def test_sincos_near_quadrant_edges():
    """Doubles within a few thousand ulps of q pi/2 (q = 1..4), where the
    reduced argument is tiny and its second word and tail decide the result,
    and of the quadrant boundaries (2q+1) pi/4, and the smallest angles."""
    xs = []
    for q in range(1, 5):
        c = np.float64(q * np.pi / 2)
        xs.append(c + np.arange(-3000, 3001) * np.spacing(c))
    for q in range(0, 8, 2):
        c = np.float64((q + 1) * np.pi / 4)
        xs.append(c + np.arange(-3000, 3001) * np.spacing(c))
    xs.append(2 * np.pi * (1.0 + np.arange(0, 1000)) * U01_SCALE)
    x = np.concatenate(xs)
    s, c = sincos_host(x)
    xl = x.astype(np.longdouble)
    es = ulp_error(s, np.sin(xl))
    ec = ulp_error(c, np.cos(xl))
    assert es.max() < 0.6, f"sin {es.max():.3f} ulp at {x[np.argmax(es)]!r}"
    assert ec.max() < 0.6, f"cos {ec.max():.3f} ulp at {x[np.argmax(ec)]!r}"


def test_sincos_symmetries():
    """sin(pi - x) and sin x differ only by the rounding of pi - x; the
    quadrant selection must give sin and cos the right signs in all four
    quadrants, and s^2 + c^2 = 1 to rounding."""
    x = 2 * np.pi * uniforms(200_000, 4)
    s, c = sincos_host(x)
    quadrant = np.floor(x / (np.pi / 2)).astype(int) % 4
    assert np.all(np.sign(s)[quadrant < 2] >= 0) and np.all(np.sign(s)[quadrant >= 2] <= 0)
    assert np.all(np.sign(c)[(quadrant == 0) | (quadrant == 3)] >= 0)
    assert np.all(np.sign(c)[(quadrant == 1) | (quadrant == 2)] <= 0)
    assert np.max(np.abs(s * s + c * c - 1)) < 4e-16


def test_box_muller_pairs_against_the_oracle():
    """The pair sqrt(-2 log u0) (cos, sin)(2 pi u1) from the host build
    against the oracle's (host libm): equal to the last bit for almost every
    pair, within 2 ulp for all."""
    from oracle.oracle import Oracle
    kura = Oracle("kuramoto")
    dt = 1.0
    n = 4000
    got = np.empty((n, 2))
    want = np.empty((n, 2))
    for i in range(n):
        w0, w1 = kura.threefry((i, 12345 + i), (0, 0xF00DCAFE))
        u0, u1 = kura.u01(w0), kura.u01(w1)
        r = np.sqrt(-2 * log_host(np.array([u0]))[0])
        s, c = sincos_host(np.array([2 * np.pi * u1]))
        got[i] = (r * c[0], r * s[0])
        want[i] = kura.noise_pair(i, 12345 + i, dt)
    ulps = np.abs(got - want) / np.spacing(np.abs(want))
    assert ulps.max() <= 2.0
    assert np.mean(got == want) > 0.97


def _planted(tmp_path, old, new):
    src = open(os.path.join(DEVICE, "gauss.hip")).read()
    assert src.count(old) == 1, old
    d = tmp_path / "dev"
    d.mkdir()
    (d / "gauss.hip").write_text(src.replace(old, new))
    return build(str(d), str(tmp_path))


def test_dropping_the_reduction_tail_fails_the_bound(tmp_path):
    """The sine needs the reduction's rounding error t wherever cos r t is
    a sizable part of an ulp of sin r; the same source without it exceeds
    the bound."""
    lib = _planted(tmp_path, "ws = fma(t, w, ws);", "")
    x = 2 * np.pi * uniforms(400_000, 6)
    s, _ = sincos_host(x, lib)
    assert ulp_error(s, np.sin(x.astype(np.longdouble))).max() > 0.6


def test_dropping_the_square_remainder_fails_the_bound(tmp_path):
    """The logarithm's f^2/2 carries its rounding error; the same source
    without it exceeds the bound."""
    lib = _planted(tmp_path, "lo - hfsq_l", "lo")
    u = uniforms(400_000, 5)
    assert ulp_error(log_host(u, lib), np.log(u.astype(np.longdouble))).max() > 0.66


def test_constants_match_derivation():
    """The constants in gauss.hip are those gauss_coefficients.py derives
    (mpmath at 60 digits)."""
    pytest.importorskip("mpmath")
    out = subprocess.run([sys.executable, os.path.join(DEVICE, "gauss_coefficients.py")],
                         check=True, capture_output=True, text=True).stdout
    src = open(os.path.join(DEVICE, "gauss.hip")).read()
    for name in ("sqrt\\(1/2\\)", "L1", "L2", "2/pi", "P1", "P2"):
        value = re.search(rf"^{name}\s+(\S+)", out, re.M).group(1)
        assert value in src, name
    for name in ("G", "S", "C"):
        for q in re.search(rf"^{name}\s+\[(.*)\]", out, re.M).group(1).replace("'", "").split(", "):
            assert q in src, (name, q)
    m = re.search(r"^-1/6\s+(\S+)\s+remainder\s+(\S+)", out, re.M)
    assert m.group(1) in src and m.group(2) in src
