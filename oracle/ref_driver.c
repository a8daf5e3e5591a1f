/*
 * ref_driver.c -- builds the reference's own OpenCL C kernels as host C.
 *
 * TEST INFRASTRUCTURE ONLY (oracle/_ref).  It is compiled only in a tree
 * where the reference sources are present (see oracle/Makefile); the
 * resulting libraries validate oracle/sodecl_oracle.c and serve as the
 * CPU baseline in bench.py.
 *
 * The reference assembles one OpenCL program per call as
 *   [Random123 includes] + #defines + system source + solver + caller
 * (src/sodeclmgr.cpp:374-503).  This driver includes the same files in the
 * same order.  OpenCL C differs from C99 here only in three keywords, which
 * are mapped below: __kernel and __global (no meaning on a host) and
 * get_global_id(0) (the orbit the host loop is currently integrating).  The
 * compile-time macros become run-time globals; every fixed-size array of
 * the kernels then is a C99 variable-length array, arithmetic unchanged.
 *
 * Build-time macros (set by oracle/Makefile):
 *   REF_SYSTEM_FILE  the system plugin source
 *   REF_SOLVER_FILE  src/kernels/{euler,rk4,ie,im,stochastic_euler}.cl
 *   REF_CALLER_FILE  src/kernels/{solver_caller,stochastic_solver_caller}.cl
 *   REF_STOCHASTIC   1 when the caller is the stochastic one
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int ref_numeq = 1, ref_numpar = 1, ref_numnoi = 0, ref_listsize = 1, ref_numsteps = 1;
double ref_kdt = 0.0;
double ref_epsilon1 = 1e-07;
static __thread int ref_gid;

#define _numeq_ ref_numeq
#define _numpar_ ref_numpar
#define _numnoi_ ref_numnoi
#define _listsize_ ref_listsize
#define _numsteps_ ref_numsteps
#define _m_dt_ ref_kdt
#define _epsilon1_ ref_epsilon1

#define __kernel
#define __global
#define get_global_id(dim) ref_gid

#if REF_STOCHASTIC
#define R123_USE_U01_DOUBLE 1
#include <Random123/threefry.h>
#include <Random123/u01.h>
#endif

#include REF_SYSTEM_FILE
#include REF_SOLVER_FILE
#include REF_CALLER_FILE

/* Same rule as sodeclmgr.cpp:951 (see oracle/sodecl_oracle.c). */
static long ref_num_launches(double tspan, double dt, int ksteps)
{
    double x = tspan / (dt * ksteps);
    if (!(x > 0)) return 0;
    if (x > 2147483647.0) return -1;
    long n = 0;
    n = (long)ceil(x);
    while (n > 0 && !((double)(n - 1) < x)) n--;
    while ((double)n < x) n++;
    return n;
}

/*
 * Host loop of sodeclmgr::run_sode_solver (src/sodeclmgr.cpp:921-1035):
 * before every launch the state is read back and appended to the output in
 * [equation][orbit] order; one more snapshot after the last launch.  Each
 * launch runs the caller kernel for every orbit (in parallel over orbits).
 *   t0[orbits], y0[orbits*neq] (AoS), params[orbits*npar], counters[orbits]
 *   (double, SDE only), out[nsnap][neq][orbits] or NULL.
 * Returns the number of launches, or -1.
 */
long ref_run(int neq, int npar, int nnoi, int orbits, int ksteps, double dt, double tspan,
             double* t0, double* y0, double* params, double* counters, double* out, int nthreads)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%g", dt);
    ref_kdt = strtod(buf, NULL);
    ref_numeq = neq;
    ref_numpar = npar;
    ref_numnoi = nnoi;
    ref_listsize = orbits;
    ref_numsteps = ksteps;
    long launches = ref_num_launches(tspan, dt, ksteps);
    if (launches < 0) return -1;
#ifdef _OPENMP
    int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    long slab = (long)neq * orbits;
    for (long j = 0; j <= launches; j++) {
        if (out) {
            for (int e = 0; e < neq; e++)
                for (int o = 0; o < orbits; o++)
                    out[j * slab + (long)e * orbits + o] = y0[(long)o * neq + e];
        }
        if (j == launches) break;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
        for (int o = 0; o < orbits; o++) {
            ref_gid = o;
#if REF_STOCHASTIC
            solver_caller(t0, y0, params, counters);
#else
            (void)counters;
            solver_caller(t0, y0, params);
#endif
        }
    }
    return launches;
}
