// Device helpers visible to the system plugin (emitted before it).
//
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
// Butterfly over the group: at every level the two partners add the same two
// numbers, so every lane ends with bit-identical sums.
__device__ inline double sodecl_group_sum(double v)
{
#pragma unroll
    for (int o = SODECL_GROUP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SODECL_GROUP);
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
// (no LDS traffic), cross-lane permute otherwise.  CTRL is the quad_perm
// selector: lane q of each quad reads lane q ^ m.
template <int CTRL>
__device__ inline double sodecl_quad_perm(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ inline double sodecl_group_xor(double v, int m)
{
    switch (m) {
    case 1: return sodecl_quad_perm<0xB1>(v);
    case 2: return sodecl_quad_perm<0x4E>(v);
    case 3: return sodecl_quad_perm<0x1B>(v);
    default: return __shfl_xor(v, m, SODECL_GROUP);
    }
}
#define SODECL_TEAM_API 1
#endif
