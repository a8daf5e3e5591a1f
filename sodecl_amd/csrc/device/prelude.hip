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
