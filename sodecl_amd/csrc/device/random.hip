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
SODECL_INLINE sodecl_u64 sodecl_rotl(sodecl_u64 v, int n) { return (v << n) | (v >> (64 - n)); }

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

