"""Derivation of the constants in sin.hip (run: python sin_coefficients.py).

pi and 1/pi rounded to double, the remainder pi - P1 rounded, -1/6 as a sum
of two doubles, and the polynomial P (degree 7 in z = r^2) approximating
(sin r - r + r^3/6) / r^5 on |r| <= pi/2 (slightly widened), by Chebyshev
interpolation at 60 significant digits (mpmath.chebyfit, a near-minimax
fit), with its maximum error.  The printed hex literals are the ones the
source uses, highest power first."""
import mpmath as mp

mp.mp.dps = 60


def as_double(v):
    return float(mp.mpf(v))


def remainder_over_r5(z):
    """(sin r - r + r^3/6) / r^5 as a function of z = r^2."""
    if z == 0:
        return mp.mpf(1) / 120
    r = mp.sqrt(z)
    return (mp.sin(r) - r + r ** 3 / 6) / r ** 5


def main():
    pi = mp.pi
    p1 = as_double(pi)
    p2 = as_double(pi - p1)
    print("1/pi ", float.hex(as_double(1 / pi)))
    print("P1   ", float.hex(p1))
    print("P2   ", float.hex(p2), " |pi - P1 - P2| =", mp.nstr(abs(pi - p1 - p2), 3))
    s1 = as_double(mp.mpf(-1) / 6)
    print("s1   ", float.hex(s1), " s1l", float.hex(as_double(mp.mpf(-1) / 6 - s1)))
    zmax = (pi / 2) ** 2 * mp.mpf("1.0001")
    poly, err = mp.chebyfit(remainder_over_r5, [0, zmax], 8, error=True)
    print("P    ", [float.hex(as_double(c)) for c in poly])
    print("max |P - remainder| =", mp.nstr(err, 3), "; times (pi/2)^5:", mp.nstr(err * (pi / 2) ** 5, 3))


if __name__ == "__main__":
    main()
