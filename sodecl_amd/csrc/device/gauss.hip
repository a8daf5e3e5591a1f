// The logarithm, sine and cosine of the SDE noise's Box-Muller transform
// (stochastic_solver_caller.cl:54-65: sqrt(-2 log u0) cos(2 pi u1) and
// sqrt(-2 log u0) sin(2 pi u1)), for the arguments the open-open uniforms
// give: u0 in [2^-53, 1) and x = 2 pi u1 in (0, 2 pi].  Neither needs the
// general functions' handling of special values, large arguments or
// subnormals, so each is a short straight-line sequence in place of the
// device library's (whose logarithm works in double-double throughout).
// Both are IEEE double arithmetic (fma, products, sums, one division,
// frexp and ldexp) with contraction off, so the file is also valid host
// C++ (define SODECL_MATH_FN first) and its host build computes the same
// bits as the GPU; gauss_coefficients.py derives every constant.
// Measured against x87 long double on 10^7 arguments each: at most 0.66 ulp
// for the logarithm of the uniforms, 0.58 ulp for the sine and 0.57 ulp for
// the cosine on (0, 2 pi]; correctly rounded in more than 99% of cases.

#ifndef SODECL_MATH_FN
#define SODECL_MATH_FN __host__ __device__ inline
#endif

#if defined(__clang__)
#define SODECL_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define SODECL_NO_CONTRACT
#endif

// log u for a positive normal double u.
// Method: u = 2^k m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), and
// log m = 2 atanh s with s = f / (2 + f), |s| <= 3 - 2 sqrt 2 (Cody and
// Waite, "Software Manual for the Elementary Functions", 1980): since
// 2s = f - s f,
//   log m = f - f^2/2 + s (f^2/2 + R(z)),  R(z) = 2 (atanh s - s) / s = z G(z),
// z = s^2, G of degree 7 within 2.1e-18 of R/z (3.1e-20 relative in log m).
// f - f^2/2 is kept as a sum of two doubles (the product's rounding error by
// a fused multiply-add, the difference's by Fast2Sum, |f| >= f^2), and
// k ln2 = k L1 + k L2 with k L1 exact and its sum with the leading part
// again split exactly, so only the last addition rounds at full weight.
SODECL_MATH_FN double sodecl_log_positive(double u)
{
    SODECL_NO_CONTRACT
    int e;
    const double m0 = frexp(u, &e);                       // u = m0 2^e, m0 in [1/2, 1)
    const int low = m0 < 0x1.6a09e667f3bcdp-1 ? 1 : 0;    // m0 < sqrt(1/2)
    const double m = ldexp(m0, low);
    const double dk = (double)(e - low);                  // u = 2^k m
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double g = 0x1.0c0aff044a970p-3;
    g = sodecl_fma_k(g, z, 0x1.0fbde0f4ad17bp-3);
    g = sodecl_fma_k(g, z, 0x1.3b1c3c1c81c8fp-3);
    g = sodecl_fma_k(g, z, 0x1.745cf8e4bba1bp-3);
    g = sodecl_fma_k(g, z, 0x1.c71c7201a55d7p-3);
    g = sodecl_fma_k(g, z, 0x1.2492492476a1ap-2);
    g = sodecl_fma_k(g, z, 0x1.9999999999a39p-2);
    g = sodecl_fma_k(g, z, 0x1.5555555555555p-1);
    const double hf = 0.5 * f;
    const double hfsq = hf * f;
    const double hfsq_l = fma(hf, f, -hfsq);              // f^2/2 = hfsq + hfsq_l
    const double hi = f - hfsq;
    const double lo = (f - hi) - hfsq;                    // f - hfsq = hi + lo
    const double c = fma(s, fma(z, g, hfsq), lo - hfsq_l);   // log m = hi + c
    const double kl = dk * 0x1.62e42fee00000p-1;          // k L1, exact (L1 has 32 bits)
    const double a = kl + hi;
    const double err = (kl - a) + hi;                     // |kl| >= |hi| unless k = 0, where a = hi
    return a + fma(dk, 0x1.a39ef35793c76p-33, err + c);
}

// sin x and cos x for 0 <= x <= 2 pi (and a little beyond).
// Method: q = round(x 2/pi) by one fused multiply-add with 1.5 2^52 (the
// sum's low word is q), and r = x - q pi/2 by Cody and Waite's additive
// reduction with pi/2 = P1 + P2 (|pi/2 - P1 - P2| < 1.5e-33): r1 = x - q P1
// is exact (|r1| < 1 is a multiple of 2^-53), r = r1 - q P2 rounds and
// t = (r1 - r) - q P2 is that rounding, so r + t is x - q pi/2 to 1e-32.
// On |r| <= pi/4 (z = r^2):
//   sin(r + t) = r + r^3 (-1/6 + z S(z)) + t (1 - z/2),
//   cos(r + t) = 1 - z/2 + z^2 C(z) - t r,
// S and C of degree 5 within 3.2e-20 and 6.9e-19 relative.  The sine's
// -r^3/6 (up to 0.08 of the result) is formed without error: r^3 = v + vl
// from exact products, -1/6 as a sum of two doubles, and u = r - v/6 with
// its rounding error (r - u is exact); the cosine's 1 - z/2 keeps its
// rounding error (Fast2Sum) and z's (r^2 = z + zl).  Then by the quadrant
// q mod 4: sin x = +-sin r or +-cos r, cos x = +-cos r or +-sin r.
SODECL_MATH_FN void sodecl_sincos_reduced(double x, double* sx, double* cx)
{
    SODECL_NO_CONTRACT
    const double kk = fma(x, 0x1.45f306dc9c883p-1, 0x1.8p52);
    const double q = kk - 0x1.8p52;
    const double r1 = fma(-q, 0x1.921fb54442d18p+0, x);
    const double r = fma(-q, 0x1.1a62633145c07p-54, r1);
    const double t = fma(-q, 0x1.1a62633145c07p-54, r1 - r);
    const double z = r * r;
    const double zl = fma(r, r, -z);                      // r^2 = z + zl
    // sine
    const double v = z * r;
    const double vl = fma(zl, r, fma(z, r, -v));          // r^3 = v + vl
    double sp = -0x1.ab933c1a64520p-41;
    sp = sodecl_fma_k(sp, z, 0x1.61225a3153150p-33);
    sp = sodecl_fma_k(sp, z, -0x1.ae6455332472dp-26);
    sp = sodecl_fma_k(sp, z, 0x1.71de3a550c541p-19);
    sp = sodecl_fma_k(sp, z, -0x1.a01a01a019ed4p-13);
    sp = sodecl_fma_k(sp, z, 0x1.1111111111111p-7);
    const double s1 = -0x1.5555555555555p-3;              // -1/6 = s1 + s1l
    const double u = fma(v, s1, r);
    const double ue = fma(v, s1, r - u);                  // r + v s1 = u + ue
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    double ws = fma(v, sodecl_fma_k(z, sp, -0x1.5555555555555p-57), ue);
    ws = fma(vl, s1, ws);
    ws = fma(t, w, ws);
    const double sin_r = u + ws;
    // cosine
    const double ce = (1.0 - w) - hz;                     // 1 - hz = w + ce
    double cp = -0x1.907cd8ff5ede7p-37;
    cp = sodecl_fma_k(cp, z, 0x1.1eeb67d026ec6p-29);
    cp = sodecl_fma_k(cp, z, -0x1.27e4fa16cc2d1p-22);
    cp = sodecl_fma_k(cp, z, 0x1.a01a019f4ddfdp-16);
    cp = sodecl_fma_k(cp, z, -0x1.6c16c16c16963p-10);
    cp = sodecl_fma_k(cp, z, 0x1.5555555555555p-5);
    const double z2 = z * z;
    double cb = fma(-0.5, zl, ce);
    cb = fma(-t, r, cb);
    cb = fma(z2, cp, cb);
    const double cos_r = w + cb;
    // quadrant
    unsigned n;   // low word of kk (little-endian on the host and the GPU): q mod 2^32
    __builtin_memcpy(&n, &kk, sizeof n);
    const double a = (n & 1u) ? cos_r : sin_r;
    const double b = (n & 1u) ? sin_r : cos_r;
    *sx = sodecl_with_hi_word(a, sodecl_hi_word(a) ^ ((n & 2u) << 30));
    *cx = sodecl_with_hi_word(b, sodecl_hi_word(b) ^ (((n + 1u) & 2u) << 30));
}
