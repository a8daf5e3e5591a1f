// Batched integrator, one kernel per system-size class (emitted once per
// class, after that class's namespace, with SODECL_BATCH_KERNEL the kernel's
// name, SODECL_BATCH_ITEM the class's item routine and SODECL_BATCH_WAVES
// its occupancy hint).  Workgroup k (one wavefront) integrates work item k
// of the class's list, which the host orders longest first from neq * steps:
// the dispatcher starts workgroups in index order as room frees up, so the
// costliest items start first, and the classes' kernels, launched side by
// side, share the GPU as their workgroups come and go.  An item holds one
// orbit, or up to 64 / lanes-per-orbit orbits when the class runs 32 or 16
// lanes per orbit.  A kernel per class gives each class its own register
// allocation and occupancy.
//
// Packed layouts: items[4 * item + k] = orbit or -1; the initial state of
// orbit o at y0[yoff[o] .. + neq_o]; its snapshots [snap][neq_o] at
// out[ooff[o] ...], steps[o] / ksteps + 1 of them; params[orbits][npar];
// newton[o] receives the Newton updates taken.

static_assert(SODECL_BG == 64, "one wavefront per block: the items synchronise at wavefront scope");

extern "C" __global__ void __launch_bounds__(SODECL_BG) __attribute__((amdgpu_waves_per_eu(SODECL_BATCH_WAVES)))
SODECL_BATCH_KERNEL(const int* __restrict__ items, const int* __restrict__ steps, const long long* __restrict__ yoff,
                    const long long* __restrict__ ooff, const double* __restrict__ y0,
                    const double* __restrict__ params, double* __restrict__ out, long long* __restrict__ newton)
{
    SODECL_EXP_TABLE_INIT();
    extern __shared__ double sodecl_lds[];
    SODECL_BATCH_ITEM(items + (size_t)blockIdx.x * 4, steps, yoff, ooff, y0, params, out, newton, sodecl_lds);
}
