// Counter-based random numbers shared by the integrators (compiled at run
// time by hipRTC, emitted before integrators.hip) and the ahead-of-time
// sampling kernel (static_kernels.hip).

#ifndef SODECL_INLINE
#define SODECL_INLINE __device__ __attribute__((always_inline)) inline
#endif

typedef unsigned long long sodecl_u64;

// ---------------------------------------------------------------------------
// Threefry-2x64, 20 rounds (Salmon, Moraes, Dror, Shaw, "Parallel Random
// Numbers: As Easy as 1, 2, 3", SC'11): Threefish-style add/rotate/xor rounds
// with the Skein 2x64 rotation schedule and a key injection after every four
// rounds.  Identical output to Random123's threefry2x64 (threefry.h:178-293),
// which the reference's SDE caller uses.
// ---------------------------------------------------------------------------
// v rotated left by n, 0 < n < 64 (a constant at every call).  On the GPU
// each 32-bit half of the result is one funnel shift of the two input halves
// (v_alignbit_b32: the low word of {a, b} >> s); the compiler's own lowering
// of the 64-bit shifts and or takes three to four instructions.
SODECL_INLINE sodecl_u64 sodecl_rotl(sodecl_u64 v, int n)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    unsigned rh, rl;
    if (n == 32) {
        rh = lo;
        rl = hi;
    } else if (n < 32) {
        rh = __builtin_amdgcn_alignbit(hi, lo, 32 - n);   // (hi << n) | (lo >> (32 - n))
        rl = __builtin_amdgcn_alignbit(lo, hi, 32 - n);   // (lo << n) | (hi >> (32 - n))
    } else {
        rh = __builtin_amdgcn_alignbit(lo, hi, 64 - n);   // the halves swapped, then by n - 32
        rl = __builtin_amdgcn_alignbit(hi, lo, 64 - n);
    }
    return ((sodecl_u64)rh << 32) | rl;
#else
    return (v << n) | (v >> (64 - n));
#endif
}

#define SODECL_TF_ROUND(R) \
    x0 += x1;              \
    x1 = sodecl_rotl(x1, R); \
    x1 ^= x0;

SODECL_INLINE void sodecl_threefry2x64_20(sodecl_u64 c0, sodecl_u64 c1, sodecl_u64 k0, sodecl_u64 k1,
                                          sodecl_u64* o0, sodecl_u64* o1)
{
    const sodecl_u64 k2 = 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1;
    sodecl_u64 x0 = c0 + k0;
    sodecl_u64 x1 = c1 + k1;
    SODECL_TF_ROUND(16) SODECL_TF_ROUND(42) SODECL_TF_ROUND(12) SODECL_TF_ROUND(31)
    x0 += k1; x1 += k2 + 1ull;
    SODECL_TF_ROUND(16) SODECL_TF_ROUND(32) SODECL_TF_ROUND(24) SODECL_TF_ROUND(21)
    x0 += k2; x1 += k0 + 2ull;
    SODECL_TF_ROUND(16) SODECL_TF_ROUND(42) SODECL_TF_ROUND(12) SODECL_TF_ROUND(31)
    x0 += k0; x1 += k1 + 3ull;
    SODECL_TF_ROUND(16) SODECL_TF_ROUND(32) SODECL_TF_ROUND(24) SODECL_TF_ROUND(21)
    x0 += k1; x1 += k2 + 4ull;
    SODECL_TF_ROUND(16) SODECL_TF_ROUND(42) SODECL_TF_ROUND(12) SODECL_TF_ROUND(31)
    x0 += k2; x1 += k0 + 5ull;
    *o0 = x0;
    *o1 = x1;
}

// 64-bit word -> double in the open interval (0, 1): (1 + (w >> 11)) * (2^53-1) * 2^-106,
// the open-open 64->53 conversion of Random123's u01.h:129-131.
SODECL_INLINE double sodecl_u01(sodecl_u64 w)
{
    return (double)(1ull + (w >> 11)) * 0x1.fffffffffffffp-54;
}

