"""Derivation of the constants in gauss.hip (run: python gauss_coefficients.py).

Logarithm (reduction u = 2^k m, m in [sqrt(1/2), sqrt(2)), f = m - 1,
s = f / (2 + f), log m = 2 atanh s):
  * ln 2 split as L1 + L2, L1 with its low 21 significand bits zero (k L1 is
    exact for |k| < 2^21), L2 the remainder rounded to double;
  * G of degree 7 approximating G(z) = R(z) / z, R(z) = 2 (atanh(s) - s) / s
    = sum_{k>=1} 2 z^k / (2k + 1), z = s^2, on 0 <= z <= (3 - 2 sqrt 2)^2
    (the largest s^2 of the reduction, slightly widened), by Chebyshev
    interpolation at 60 significant digits (mpmath.chebyfit, near-minimax).
Sine and cosine (reduction x = q pi/2 + r, |r| <= pi/4):
  * 2/pi rounded to double; pi/2 split as P1 + P2, P1 = pi/2 rounded to
    double, P2 the remainder rounded, with |pi/2 - P1 - P2|;
  * S of degree 5 approximating (sin r - r + r^3/6) / r^5 and C of degree 5
    approximating (cos r - 1 + r^2/2) / r^4, both as functions of z = r^2 on
    0 <= z <= (pi/4)^2 (slightly widened), by the same interpolation.
Hex literals are printed in the order the source uses them."""
import struct

import mpmath as mp

mp.mp.dps = 60


def as_double(v):
    return float(mp.mpf(v))


def hexd(v):
    return float.hex(as_double(v))


def clear_low_bits(x, bits):
    (u,) = struct.unpack("<Q", struct.pack("<d", x))
    u &= ~((1 << bits) - 1)
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def fit(fn, lo, hi, degree):
    poly, err = mp.chebyfit(fn, [lo, hi], degree + 1, error=True)
    return poly, err


def g_of_z(z):
    if z == 0:
        return mp.mpf(2) / 3
    s = mp.sqrt(z)
    return 2 * (mp.atanh(s) - s) / (s * z)


def s_of_z(z):
    if z == 0:
        return mp.mpf(1) / 120
    r = mp.sqrt(z)
    return (mp.sin(r) - r + r ** 3 / 6) / r ** 5


def c_of_z(z):
    if z == 0:
        return mp.mpf(1) / 24
    r = mp.sqrt(z)
    return (mp.cos(r) - 1 + z / 2) / z ** 2


def main():
    ln2 = mp.log(2)
    l1 = clear_low_bits(as_double(ln2), 21)
    l2 = as_double(ln2 - l1)
    print("sqrt(1/2)", hexd(mp.sqrt(mp.mpf(1) / 2)))
    print("L1       ", float.hex(l1))
    print("L2       ", float.hex(l2), " |ln2 - L1 - L2| =", mp.nstr(abs(ln2 - l1 - l2), 3))
    smax = 3 - 2 * mp.sqrt(2)
    zmax = smax ** 2 * mp.mpf("1.001")
    poly, err = fit(g_of_z, 0, zmax, 7)
    print("G        ", [hexd(c) for c in poly])
    print("max |G - R/z| =", mp.nstr(err, 3), "; in log m (times s^3, relative to 2s):", mp.nstr(err * zmax / 2, 3))

    half_pi = mp.pi / 2
    p1 = as_double(half_pi)
    p2 = as_double(half_pi - p1)
    print("2/pi     ", hexd(2 / mp.pi))
    print("P1       ", float.hex(p1))
    print("P2       ", float.hex(p2), " |pi/2 - P1 - P2| =", mp.nstr(abs(half_pi - p1 - p2), 3))
    rz = (mp.pi / 4) ** 2 * mp.mpf("1.001")
    spoly, serr = fit(s_of_z, 0, rz, 5)
    print("S        ", [hexd(c) for c in spoly])
    print("max |S - (sin r - r + r^3/6)/r^5| =", mp.nstr(serr, 3), "; times (pi/4)^5 / sin(pi/4):",
          mp.nstr(serr * (mp.pi / 4) ** 5 / mp.sin(mp.pi / 4), 3))
    cpoly, cerr = fit(c_of_z, 0, rz, 5)
    print("C        ", [hexd(c) for c in cpoly])
    print("max |C - (cos r - 1 + r^2/2)/r^4| =", mp.nstr(cerr, 3), "; times (pi/4)^4 / cos(pi/4):",
          mp.nstr(cerr * (mp.pi / 4) ** 4 / mp.cos(mp.pi / 4), 3))
    print("-1/6     ", hexd(mp.mpf(-1) / 6), " remainder", hexd(mp.mpf(-1) / 6 - as_double(mp.mpf(-1) / 6)))


if __name__ == "__main__":
    main()
