/*
 * ref_simd.c -- the reference's own OpenCL C kernels as host C with the
 * orbit loop vectorised: the CPU baseline of bench.py (oracle/_ref/libref_*_simd.so).
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY.  Built only where the reference
 * sources are present (oracle/Makefile), like ref_driver.c.
 *
 * An OpenCL CPU runtime vectorises a kernel across work-items: it packs
 * neighbouring work-items (orbits) into SIMD lanes and runs the kernel body
 * for all of them at once.  This driver does the same for the reference's
 * ODE path: the solver (REF_SOLVER_FILE, src/kernels/{euler,rk4}.cl) and the
 * system plugin (REF_SYSTEM_FILE) are included unchanged, with the sizes
 * fixed at compile time (REF_NEQ, REF_NPAR) so that their loops unroll, and
 * the caller kernel's per-work-item loop (src/kernels/solver_caller.cl:12-50:
 * load t, y and the parameters, ksteps times sode_solver then t += dt, store)
 * runs for REF_W orbits in lockstep, the orbit loop innermost and
 * vectorised (#pragma omp simd; clang vectorises it with the vector
 * transcendentals of glibc's libmvec, -fveclib=libmvec).  The host loop
 * around it is src/sodeclmgr.cpp:921-1035, as in ref_driver.c, with OpenMP
 * over blocks of orbits.
 *
 * Vector exp/sin are within 4 ulp rather than the scalar libm's 1; these
 * libraries time the reference arithmetic, they are not a parity checker.
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

#ifndef REF_W
#define REF_W 8
#endif

static int ref_numsteps = 1, ref_listsize = 1;
static double ref_kdt = 0.0;

#define _numeq_ REF_NEQ
#define _numpar_ REF_NPAR
#define _numnoi_ 0
#define _listsize_ ref_listsize
#define _numsteps_ ref_numsteps
#define _m_dt_ ref_kdt
#define _epsilon1_ 1e-07
#define __kernel
#define __global

#include REF_SYSTEM_FILE
#include REF_SOLVER_FILE

/* Same rule as sodeclmgr.cpp:951 (see oracle/sodecl_oracle.c). */
static long ref_num_launches(double tspan, double dt, int ksteps)
{
    double x = tspan / (dt * ksteps);
    if (!(x > 0)) return 0;
    if (x > 2147483647.0) return -1;
    long n = (long)ceil(x);
    while (n > 0 && !((double)(n - 1) < x)) n--;
    while ((double)n < x) n++;
    return n;
}

/* ksteps steps of REF_W orbits in lockstep; state in structure-of-arrays
 * form (Y[eq][lane]), each lane's step the caller kernel's. */
static __attribute__((flatten)) void ref_block_steps(int ksteps, double* restrict T, double (*restrict Y)[REF_W],
                            double (*restrict P)[REF_W])
{
    for (int it = 0; it < ksteps; ++it) {
#pragma omp simd
        for (int w = 0; w < REF_W; ++w) {
            double y[_numeq_], detterm[_numeq_], params[_numpar_ > 0 ? _numpar_ : 1];
            for (int m = 0; m < _numeq_; ++m) y[m] = Y[m][w];
            for (int m = 0; m < _numpar_; ++m) params[m] = P[m][w];
            sode_solver(_m_dt_, T[w], y, detterm, params);
            T[w] = T[w] + _m_dt_;
            for (int m = 0; m < _numeq_; ++m) Y[m][w] = detterm[m];
        }
    }
}

/* Same arguments and results as ref_driver.c's ref_run (deterministic
 * solvers only: nnoi must be 0 and neq/npar the compiled sizes). */
long ref_run(int neq, int npar, int nnoi, int orbits, int ksteps, double dt, double tspan, double* t0, double* y0,
             double* params, double* counters, double* out, int nthreads)
{
    (void)counters;
    if (neq != REF_NEQ || npar != REF_NPAR || nnoi != 0) return -1;
    char buf[64];
    snprintf(buf, sizeof buf, "%g", dt);
    ref_kdt = strtod(buf, NULL);
    ref_listsize = orbits;
    ref_numsteps = ksteps;
    long launches = ref_num_launches(tspan, dt, ksteps);
    if (launches < 0) return -1;
#ifdef _OPENMP
    int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    long slab = (long)neq * orbits;
    for (long j = 0; j <= launches; j++) {
        if (out) {
            for (int e = 0; e < neq; e++)
                for (int o = 0; o < orbits; o++) out[j * slab + (long)e * orbits + o] = y0[(long)o * neq + e];
        }
        if (j == launches) break;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
        for (int b = 0; b < orbits; b += REF_W) {
            double T[REF_W], Y[REF_NEQ][REF_W], P[REF_NPAR > 0 ? REF_NPAR : 1][REF_W];
            for (int w = 0; w < REF_W; ++w) {
                const int o = b + w < orbits ? b + w : orbits - 1;   /* a partial block repeats its last orbit */
                T[w] = t0[o];
                for (int m = 0; m < REF_NEQ; ++m) Y[m][w] = y0[(long)o * REF_NEQ + m];
                for (int m = 0; m < REF_NPAR; ++m) P[m][w] = params[(long)o * REF_NPAR + m];
            }
            ref_block_steps(ksteps, T, Y, P);
            for (int w = 0; w < REF_W && b + w < orbits; ++w) {
                const int o = b + w;
                t0[o] = T[w];
                for (int m = 0; m < REF_NEQ; ++m) y0[(long)o * REF_NEQ + m] = Y[m][w];
            }
        }
    }
    return launches;
}
