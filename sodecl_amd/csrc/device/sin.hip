// Sine of a double for system sources: a call sin(x) in the plugin resolves
// to sodecl_sin(x) unless the caller asks for the device library's sine.
// The file is also valid host C++ (define SODECL_MATH_FN first), so that the
// same operations can be checked on the CPU.
//
// Method: Cody and Waite's additive range reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) by pi, with the
// rounding error of the reduction carried as a tail, and a polynomial for
// the series remainder fitted by Chebyshev interpolation (near-minimax):
//   sin x = sign(x) (-1)^k sin r,  k = round(|x| / pi),  r = |x| - k pi.
//   pi ~ P1 + P2: P1 is pi rounded to double, P2 the remainder rounded
//   (|pi - P1 - P2| < 3.1e-33).  r1 = |x| - k P1 is exact: for |x| >= 2 both
//   terms are multiples of 2^-51 and |r1| < 2; for |x| < 2, k is 0 or 1 and
//   both are multiples of 2^-52.  r2 = round(r1 - k P2); while
//   |r1| >= 2 k P2, r1 - r2 is exact (Sterbenz) and t = r1 - r2 - k P2 is
//   r2's rounding error, so r = r2 + t up to k |pi - P1 - P2|, which is
//   below 2^-56 |r| whenever |r2| >= 4 k P2.
//   sin(r2 + t) = r2 - r2^3 / 6 + r2^5 P(r2^2) + t (1 - z/2 + z^2/24),
//   z = r2^2, P of degree 7 in z, within 7.7e-21 of (sin r - r + r^3/6) / r^5
//   on |r| <= pi/2 (sin_coefficients.py derives every constant here); the tail's omitted terms
//   are below 0.02 ulp.  The leading correction -r2^3 / 6 (up to 0.65 in size
//   against a result near 1) is formed without error: r2^3 = v + vl from
//   exact products, -1/6 as a sum of two doubles, and u = r2 + v (-1/6) with
//   its rounding error e (r2 - u is exact since |v| / 6 <= |r2| / 2).
//   Measured on 2e7 arguments: at most 0.88 ulp from the exact sine.
// The device library's sin(x) is used instead where this does not hold:
// |x| > 2^30 (and NaN or infinite x), and x within 4 k P2 (about
// k 4.9e-16) of a nonzero multiple k pi.
//
// Cost on gfx950: about 35 VALU instructions, against about 54 for the
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

// fma(a, b, k) for a constant addend k, as one three-operand instruction
// that reads k from a scalar register pair.  Left to the compiler, a chain
// step whose addend is a constant became the two-operand form, whose addend
// is also its destination: a copy of the constant into vector registers
// before every step, or, hoisted out of a loop, two vector registers held
// for each constant through the whole loop (fewer wavefronts per SIMD).
// k must be the same in every lane (a literal constant at every call).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double sodecl_fma_k(double a, double b, double k)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(k));
    return d;
}
#else
#define sodecl_fma_k fma
#endif

// sin x by the reduction above; *library is set where sodecl_sin returns the
// device library's sin(x) instead.
// Every product that feeds a sum is an explicit fma, so contraction cannot
// change the result.
SODECL_MATH_FN double sodecl_sin_reduced(double x, bool* library)
{
    const double ax = fabs(x);
    const double p1 = 0x1.921fb54442d18p+1;
    const double p2 = 0x1.1a62633145c07p-53;
    // kk = k + 2^52 with k = round(|x| / pi): k's parity is kk's lowest bit
    const double kk = fma(ax, 0x1.45f306dc9c883p-2, 0x1p52);
    const double k = kk - 0x1p52;
    const double r1 = fma(-k, p1, ax);
    const double r2 = fma(-k, p2, r1);
    const double t = fma(-k, p2, r1 - r2);
    *library = !(ax <= 0x1p30) || fabs(r2) < k * (4.0 * p2);
    const double z = r2 * r2;
    double p = -0x1.283dfd82ab975p-57;
    p = sodecl_fma_k(p, z, 0x1.951051ba25afep-49);
    p = sodecl_fma_k(p, z, -0x1.ae7f03ca459a2p-41);
    p = sodecl_fma_k(p, z, 0x1.612460f6bebccp-33);
    p = sodecl_fma_k(p, z, -0x1.ae64567eff356p-26);
    p = sodecl_fma_k(p, z, 0x1.71de3a556c3ecp-19);
    p = sodecl_fma_k(p, z, -0x1.a01a01a01a018p-13);
    p = sodecl_fma_k(p, z, 0x1.1111111111111p-7);
    const double zl = fma(r2, r2, -z);       // r2^2 = z + zl
    const double v = z * r2;                 // r2^3 = v + vl
    const double vl = fma(zl, r2, fma(z, r2, -v));
    const double s1 = -0x1.5555555555555p-3;     // -1/6 = s1 + s1l
    const double s1l = -0x1.5555555555555p-57;
    const double u = fma(v, s1, r2);
    const double e = fma(v, s1, r2 - u);     // r2 + v s1 = u + e
    double w = fma(v, sodecl_fma_k(z, p, s1l), e);    // + v (s1l + z p)
    w = fma(vl, s1, w);
    w = fma(t, fma(z, fma(z, 0x1.5555555555555p-5, -0.5), 1.0), w);   // + t (1 - z/2 + z^2/24)
    const double s = u + w;
    // sin x = sign(x) (-1)^k sin r
    unsigned lo;   // low word of kk (little-endian on the host and the GPU): k mod 2^32
    __builtin_memcpy(&lo, &kk, sizeof lo);
    const unsigned sx = sodecl_hi_word(x) & 0x80000000u;
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned hi = __builtin_amdgcn_bitop3_b32(sodecl_hi_word(s), lo << 31, sx, 0x96);   // a ^ b ^ c
#else
    const unsigned hi = sodecl_hi_word(s) ^ (lo << 31) ^ sx;
#endif
    return sodecl_with_hi_word(s, hi);
}

SODECL_MATH_FN double sodecl_sin(double x)
{
    bool library;
    const double s = sodecl_sin_reduced(x, &library);
    if (__builtin_expect(library, 0)) return sin(x);
    return s;
}

// Single precision keeps the library's sinf; integer arguments convert to
// double, as the library's sin(int) does.
SODECL_MATH_FN float sodecl_sin(float x) { return sinf(x); }
template <typename T>
SODECL_MATH_FN double sodecl_sin(T x) { return sodecl_sin((double)x); }
