"""Derivation of the constants in exp.hip (run: python exp_coefficients.py,
or with --table to print the table in the form the source embeds).

Tang's table-driven method with N = 256 table points:
  * 2^(j/256) for j = 0..255, each as a double T_hi (rounded to nearest) and
    the remainder T_lo = 2^(j/256) - T_hi rounded to double, at 60
    significant digits;
  * 256/ln 2 rounded to double; ln 2/256 split as C1 + C2, C1 with its low
    22 significand bits zero (so n C1 is exact for |n| < 2^22), C2 the
    remainder rounded to double;
  * Q of degree 2 approximating (exp(r) - 1 - r) / r^2 on |r| <= ln 2/512
    (slightly widened), by Chebyshev interpolation at 60 significant digits
    (mpmath.chebyfit, a near-minimax fit), with its maximum error.
Hex literals are printed in the order the source uses them."""
import struct
import sys

import mpmath as mp

mp.mp.dps = 60
N = 256


def as_double(v):
    return float(mp.mpf(v))


def clear_low_bits(x, bits):
    (u,) = struct.unpack("<Q", struct.pack("<d", x))
    u &= ~((1 << bits) - 1)
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def q_of_r(r):
    if r == 0:
        return mp.mpf(1) / 2
    return (mp.expm1(r) - r) / r ** 2


def table():
    out = []
    for j in range(N):
        v = mp.power(2, mp.mpf(j) / N)
        hi = as_double(v)
        lo = as_double(v - hi)
        out.append((hi, lo))
    return out


def main():
    ln2 = mp.log(2)
    print("256/ln2 ", float.hex(as_double(N / ln2)))
    c1 = clear_low_bits(as_double(ln2 / N), 22)
    c2 = as_double(ln2 / N - c1)
    print("C1      ", float.hex(c1))
    print("C2      ", float.hex(c2), " |ln2/256 - C1 - C2| =", mp.nstr(abs(ln2 / N - c1 - c2), 3))
    rmax = ln2 / (2 * N) * mp.mpf("1.001")
    poly, err = mp.chebyfit(q_of_r, [-rmax, rmax], 3, error=True)
    print("Q       ", [float.hex(as_double(c)) for c in poly])
    print("max |Q - (exp(r)-1-r)/r^2| =", mp.nstr(err, 3), "; times rmax^2:", mp.nstr(err * rmax ** 2, 3))
    if "--table" in sys.argv:
        rows = table()
        for j in range(0, N, 2):
            a, b = rows[j], rows[j + 1]
            print(f"    {float.hex(a[0])}, {float.hex(a[1])}, {float.hex(b[0])}, {float.hex(b[1])},")


if __name__ == "__main__":
    main()
