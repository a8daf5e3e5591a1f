// Sine of a double for system sources: a call sin(x) in the plugin resolves
// to sodecl_sin(x) unless the caller asks for the device library's sine.
// The file is also valid host C++ (define SODECL_MATH_FN first), so that the
// same operations can be checked on the CPU.
//
// Method: Cody and Waite's additive range reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) by pi, with the
// rounding error of the reduction carried as a tail, and the Taylor series
// of sin:
//   k = rint(x / pi), r = x - k pi in [-pi/2, pi/2], sin x = (-1)^k sin r.
//   pi ~ P1 + P2 + P3: P1 is pi rounded to double, P2 the remainder rounded,
//   P3 the next remainder rounded (|pi - P1 - P2 - P3| < 1.2e-49).
//   r1 = x - k P1 is exact: for |x| >= 2 both x and k P1 are multiples of
//   2^-51 and |r1| < 2; for |x| < 2, k is 0 or +-1 and both are multiples of
//   2^-52.  r2 = round(r1 - k P2); while |r1| >= 2 |k P2|, r1 - r2 is exact
//   (Sterbenz) and t = (r1 - r2 - k P2) - k P3 is r2's rounding error less
//   the third term, so r = r2 + t to about 2^-100 relative.
//   sin(r2 + t) = r2 + S1 r2^3 + r2^5 (S2 + z S3 + ... + z^8 S10)
//                 + t (1 - z/2 + z^2/24),   z = r2^2, S_n = (-1)^n / (2n+1)!
//   The first omitted series term is below 1.2e-18 on |r| <= pi/2 and the
//   tail's below 0.02 ulp.  The leading correction S1 r2^3 (up to 0.65 in
//   size against a result near 1) is formed without error: r2^3 = v + vl from
//   exact products, S1 = -1/6 as a sum of two doubles, and u = r2 + v S1
//   with its rounding error e (r2 - u is exact since |v S1| <= |r2| / 2).
//   Measured on 8e6 arguments: at most 0.87 ulp from the exact sine.
// The device library's sin(x) is used instead where this does not hold:
// |x| > 2^30 (and NaN or infinite x), and x within 4 |k P2| (about
// |k| 4.9e-16) of a nonzero multiple of pi.
//
// Cost on gfx950: about 38 VALU instructions, against about 64 for the
// device library's sine (which reduces by pi/2 in double-double and evaluates
// both the sine and cosine polynomials).

#ifndef SODECL_MATH_FN
#define SODECL_MATH_FN __host__ __device__ inline
#endif

// High 32-bit word of a double, and a double with its high word replaced.
SODECL_MATH_FN unsigned sodecl_hi_word(double v)
{
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof b);
    return (unsigned)(b >> 32);
}
SODECL_MATH_FN double sodecl_with_hi_word(double v, unsigned hi)
{
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof b);
    b = (b & 0xffffffffull) | ((unsigned long long)hi << 32);
    __builtin_memcpy(&v, &b, sizeof b);
    return v;
}

// sin x by the reduction above; *library is set where sodecl_sin returns the
// device library's sin(x) instead.
// Every product that feeds a sum is an explicit fma, so contraction cannot
// change the result.
SODECL_MATH_FN double sodecl_sin_reduced(double x, bool* library)
{
    const double inv_pi = 0x1.45f306dc9c883p-2;
    const double p1 = 0x1.921fb54442d18p+1;
    const double p2 = 0x1.1a62633145c07p-53;
    const double p3 = -0x1.f1976b7ed8fbcp-109;
    const double k = rint(x * inv_pi);
    const double r1 = fma(-k, p1, x);
    const double r2 = fma(-k, p2, r1);
    double t = fma(-k, p2, r1 - r2);
    t = fma(-k, p3, t);
    *library = !(fabs(x) <= 0x1p30) || fabs(r2) < fabs(k) * (4.0 * p2);
    const double z = r2 * r2;
    double p = 0x1.71b8ef6dcf572p-66;        //  1/21!
    p = fma(p, z, -0x1.2f49b46814157p-57);   // -1/19!
    p = fma(p, z, 0x1.952c77030ad4ap-49);    //  1/17!
    p = fma(p, z, -0x1.ae7f3e733b81fp-41);   // -1/15!
    p = fma(p, z, 0x1.6124613a86d09p-33);    //  1/13!
    p = fma(p, z, -0x1.ae64567f544e4p-26);   // -1/11!
    p = fma(p, z, 0x1.71de3a556c734p-19);    //  1/9!
    p = fma(p, z, -0x1.a01a01a01a01ap-13);   // -1/7!
    p = fma(p, z, 0x1.1111111111111p-7);     //  1/5!
    const double zl = fma(r2, r2, -z);       // r2^2 = z + zl
    const double v = z * r2;                 // r2^3 = v + vl
    const double vl = fma(zl, r2, fma(z, r2, -v));
    const double s1 = -0x1.5555555555555p-3;     // -1/6 = s1 + s1l
    const double s1l = -0x1.5555555555555p-57;
    const double u = fma(v, s1, r2);
    const double e = fma(v, s1, r2 - u);     // r2 + v s1 = u + e
    double w = fma(v, fma(z, p, s1l), e);    // + v (s1l + z p)
    w = fma(vl, s1, w);
    w = fma(t, fma(z, fma(z, 0x1.5555555555555p-5, -0.5), 1.0), w);   // + t (1 - z/2 + z^2/24)
    const double s = u + w;
    // (-1)^k, and the sign of a zero argument (r2 is +0 for x = -0)
    unsigned hi = sodecl_hi_word(s) ^ ((unsigned)(int)k << 31);
    hi = x == 0.0 ? sodecl_hi_word(x) : hi;
    return sodecl_with_hi_word(s, hi);
}

SODECL_MATH_FN double sodecl_sin(double x)
{
    bool library;
    const double s = sodecl_sin_reduced(x, &library);
    if (__builtin_expect(library, 0)) return sin(x);
    return s;
}

// Single precision keeps the library's sinf.
SODECL_MATH_FN float sodecl_sin(float x) { return sinf(x); }
