// Device helpers visible to the system plugin (emitted before it).

// A double moved across lanes by one DPP control (two 32-bit moves); a lane
// whose source lies outside the wavefront receives 0.
template <int CTRL>
__device__ inline double sodecl_dpp(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
#define SODECL_DPP_XOR1 0xB1          // quad_perm [1,0,3,2]
#define SODECL_DPP_XOR2 0x4E          // quad_perm [2,3,0,1]
#define SODECL_DPP_XOR3 0x1B          // quad_perm [3,2,1,0]
#define SODECL_DPP_HALF_MIRROR 0x141  // lane i <- 7 - i within each 8 lanes
#define SODECL_DPP_MIRROR 0x140       // lane i <- 15 - i within each 16 lanes
#define SODECL_DPP_WAVE_SHR1 0x138    // lane i <- lane i - 1 across the wavefront (lane 0 <- 0)
#define SODECL_DPP_WAVE_SHL1 0x130    // lane i <- lane i + 1 across the wavefront (lane 63 <- 0)
#define SODECL_DPP_ROW_SHR(n) (0x110 + (n))   // lane i <- lane i - n within its 16-lane row (0 past the start)
#define SODECL_DPP_ROW_SHL(n) (0x100 + (n))   // lane i <- lane i + n within its 16-lane row (0 past the end)

// Orders this wavefront's LDS accesses before the call against those after
// it, for exchanges among the lanes of one wavefront: DS instructions of one
// wavefront execute in order, and the fences plus the wave barrier keep the
// compiler from moving LDS accesses across this point.  No other wavefront
// is waited for.
__device__ inline void sodecl_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Under the group schedule SODECL_GROUP consecutive lanes of a wavefront
// share one orbit.  A plugin entry point that the whole group calls together
// (sode_system_vjp_group) may use:
//   sodecl_group_size()  lanes per orbit
//   sodecl_lane()        this lane's index within the group, 0..G-1
//   sodecl_group_sum(v)  sum of v over the group, the same value on every lane
// These exist only in device code compiled for the group schedule
// (SODECL_GROUP_API is defined there).
#if SODECL_SCHEDULE == 2
__device__ inline int sodecl_group_size() { return SODECL_GROUP; }
__device__ inline int sodecl_lane() { return (int)(threadIdx.x % SODECL_GROUP); }
// Reduction over the group in levels of lane distance 1, 2, 4, 8, 16, 32.
// The first four are DPP moves inside a 16-lane row: xor 1 and xor 2 by quad
// permutations, then the half-row and row mirrors, which pair every lane with
// the opposite block of the previous level (whose lanes all hold the same
// partial sum) exactly as xor 4 and xor 8 would; the last two are cross-lane
// permutes.  At every level the two partners add the same two numbers, so
// every lane ends with a bit-identical sum.
// A group of the whole wavefront reads the four row sums into scalar
// registers instead (v_readlane) and adds them as (r0 + r1) + (r2 + r3):
// the same two levels of additions on the same operands, without the
// permutes.
__device__ inline double sodecl_readlane(double v, int lane)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ inline double sodecl_group_sum(double v)
{
    if (SODECL_GROUP >= 2) v += sodecl_dpp<SODECL_DPP_XOR1>(v);
    if (SODECL_GROUP >= 4) v += sodecl_dpp<SODECL_DPP_XOR2>(v);
    if (SODECL_GROUP >= 8) v += sodecl_dpp<SODECL_DPP_HALF_MIRROR>(v);
    if (SODECL_GROUP >= 16) v += sodecl_dpp<SODECL_DPP_MIRROR>(v);
    if (SODECL_GROUP == 64)
        return (sodecl_readlane(v, 0) + sodecl_readlane(v, 16)) + (sodecl_readlane(v, 32) + sodecl_readlane(v, 48));
    if (SODECL_GROUP >= 32) v += __shfl_xor(v, 16, SODECL_GROUP);
    return v;
}
#define SODECL_GROUP_API 1
#endif

// Under the team schedule SODECL_GROUP consecutive lanes share one orbit and
// every one of them holds the whole state.  The plugin's team entry point
// (sode_system_team), called by all lanes of the team together with the same
// arguments, may split the drift among them and must return it complete on
// every lane.  It may use:
//   sodecl_group_size()     lanes per orbit
//   sodecl_lane()           this lane's index within the team, 0..G-1
//   sodecl_group_xor(v, m)  v of team lane sodecl_lane() ^ m (m < G)
// These exist only in device code compiled for the team schedule
// (SODECL_TEAM_API is defined there).
#if SODECL_SCHEDULE == 4
__device__ inline int sodecl_group_size() { return SODECL_GROUP; }
__device__ inline int sodecl_lane() { return (int)(threadIdx.x % SODECL_GROUP); }
// Lane exchange within quads by DPP quad permutations for m = 1, 2, 3
// (no LDS traffic), cross-lane permute otherwise.
__device__ inline double sodecl_group_xor(double v, int m)
{
    switch (m) {
    case 1: return sodecl_dpp<SODECL_DPP_XOR1>(v);
    case 2: return sodecl_dpp<SODECL_DPP_XOR2>(v);
    case 3: return sodecl_dpp<SODECL_DPP_XOR3>(v);
    default: return __shfl_xor(v, m, SODECL_GROUP);
    }
}
#define SODECL_TEAM_API 1
#endif
