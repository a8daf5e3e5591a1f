// Persistent batched integrator: orbits of different system sizes and step
// counts in one launch.  Each wavefront takes the next work item from a
// device work queue (an atomic index into a longest-first item list the
// host builds from neq * steps) and integrates its orbits to the end with
// the routine of their size class (batch_class.hip, one instance per
// class); every wave exits once the queue is drained.  An item holds one
// orbit, or up to 64 / lanes-per-orbit orbits of the same class when that
// class runs 32 or 16 lanes per orbit.
//
// Packed layouts: items[4 * item + k] = orbit or -1; the initial state of
// orbit o at y0[yoff[o] .. + neq_o]; its snapshots [snap][neq_o] at
// out[ooff[o] ...], steps[o] / ksteps + 1 of them; params[orbits][npar];
// newton[o] receives the Newton updates taken.

static_assert(SODECL_BG == 64, "one wavefront per block: the items synchronise at wavefront scope");

// At least two wavefronts per SIMD (a register budget of 256 VGPRs) for these
// latency-bound wavefronts.  (At three, hipRTC gives the per-class functions
// half of 168 registers as VGPRs and moves values through AGPRs: 16% more
// VALU instructions in the 512-equation Newton loop, no faster.)
extern "C" __global__ void __launch_bounds__(SODECL_BG) __attribute__((amdgpu_waves_per_eu(2)))
sodecl_batch(const int* __restrict__ items, const int* __restrict__ cls, const int* __restrict__ steps,
             const long long* __restrict__ yoff, const long long* __restrict__ ooff, const double* __restrict__ y0,
             const double* __restrict__ params, double* __restrict__ out, long long* __restrict__ newton,
             int* __restrict__ head, int nitems)
{
    SODECL_EXP_TABLE_INIT();
    extern __shared__ double sodecl_lds[];
    for (;;) {
        int item = 0;
        if (threadIdx.x == 0) item = atomicAdd(head, 1);
        item = __shfl(item, 0, SODECL_BG);
        if (item >= nitems) break;
        const int* const io = items + (size_t)item * 4;
        switch (cls[io[0]]) {
            SODECL_BATCH_CASES
        default:
            break;
        }
        sodecl_wave_sync();   // the next item reuses the LDS rows
    }
}

// Resets the work-queue head (a kernel rather than a memset so a captured
// launch sequence replays it correctly).
extern "C" __global__ void sodecl_batch_reset(int* head)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *head = 0;
}
