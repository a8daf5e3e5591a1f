/*
 * sodecl_oracle.c -- CPU restatement of the reference SODECL integrators.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the
 * MI355X integrators: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (sodecl_amd) never calls it.
 *
 * It restates, statement by statement, the reference's device code and the
 * host segment loop (citations are to files of avramidis/sodecl):
 *   Euler                 src/kernels/euler.cl:10-18
 *   classic RK4           src/kernels/rk4.cl:17-64
 *   implicit Euler        src/kernels/ie.cl:10-117
 *   implicit midpoint     src/kernels/im.cl:3-118
 *   Euler-Maruyama        src/kernels/stochastic_euler.cl:10-22
 *   ODE caller            src/kernels/solver_caller.cl:8-51
 *   SDE caller            src/kernels/stochastic_solver_caller.cl:8-81
 *   Threefry-2x64-20      external/include/Random123/threefry.h:158-293
 *   u01 open-open 64->53  external/include/Random123/u01.h:86-131
 *   macro contract        src/sodeclmgr.cpp:374-430 (dt printed with 6
 *                         significant digits, epsilon 1e-07)
 *   segment loop + output src/sodeclmgr.cpp:921-1035
 *
 * The user system source (the plugin) is compiled into this file through
 * SODECL_SYSTEM_FILE.  The reference fixes _numeq_/_numpar_/_numnoi_ at
 * compile time; here they are run-time globals and the fixed-size arrays of
 * the reference become C99 variable-length arrays, so one build serves every
 * system size.  The arithmetic is unchanged.
 *
 * Pinning (see tests/test_oracle.py): the four oculomotor fixtures of the
 * reference's own test suite (Euler, RK4, IE, IM), Random123's published
 * Threefry-2x64-20 known-answer vectors, closed forms for u01 and the
 * Kuramoto right-hand side, and the reference kernels themselves compiled
 * as C (oracle/_ref) for every solver including Euler-Maruyama.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_numeq = 1;
int orc_numpar = 1;
int orc_numnoi = 0;
int orc_listsize = 1;
int orc_numsteps = 1;
double orc_kdt = 0.0;          /* _m_dt_: dt as printed with 6 significant digits */
double orc_epsilon1 = 1e-07;   /* _epsilon1_ (src/sodeclmgr.cpp:424-430) */
int orc_newton_maxit = 1000;   /* ie.cl:69, im.cl:65 */

#define _numeq_ orc_numeq
#define _numpar_ orc_numpar
#define _numnoi_ orc_numnoi
#define _listsize_ orc_listsize
#define _numsteps_ orc_numsteps
#define _m_dt_ orc_kdt
#define _epsilon1_ orc_epsilon1

#ifndef SODECL_SYSTEM_FILE
#error "SODECL_SYSTEM_FILE must name the system plugin source"
#endif
#include SODECL_SYSTEM_FILE

/* Optional plugin entry points: absent ones resolve to NULL. */
void sode_system_stoch(double t, double* y, double* stoch, double* p, double* noise) __attribute__((weak));
void calc_jacobian(double* y, double (*jac)[orc_numeq], double* p) __attribute__((weak));
void calc_jacobian_params(double* y, double (*jp)[orc_numpar], double* p) __attribute__((weak));
void calc_jacobian_band(double* y, double* lower, double* diag, double* upper, double* p) __attribute__((weak));
void calc_jacobian_exact(double* y, double (*jac)[orc_numeq], double* p) __attribute__((weak));

/* 1: build the Newton matrix from calc_jacobian_band when the plugin has it
 * (the same three diagonals the tridiagonal solve reads from the dense
 * Jacobian; tests check both paths give identical bits). */
int orc_use_band = 0;

enum { ORC_SEULER = 0, ORC_EULER = 1, ORC_RK4 = 2, ORC_IE = 3, ORC_IM = 4 };

/* ------------------------------------------------------------------------ */
/* Threefry-2x64 with 20 rounds (threefry.h:178-293; rotation constants     */
/* :86-93; key-schedule parity :170-171).                                   */
/* ------------------------------------------------------------------------ */
static const unsigned orc_rot64x2[8] = {16, 42, 12, 31, 16, 32, 24, 21};

static uint64_t orc_rotl64(uint64_t x, unsigned n)
{
    return (x << (n & 63)) | (x >> ((64 - n) & 63));
}

void orc_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2])
{
    uint64_t ks[3];
    ks[0] = key[0];
    ks[1] = key[1];
    ks[2] = 0x1BD11BDAA9FC1A22ULL ^ key[0] ^ key[1];
    uint64_t x0 = ctr[0] + ks[0];
    uint64_t x1 = ctr[1] + ks[1];
    for (int r = 0; r < 20; r++) {
        x0 += x1;
        x1 = orc_rotl64(x1, orc_rot64x2[r % 8]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            int s = (r + 1) / 4;      /* injection number 1..5 */
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3];
            x1 += (uint64_t)s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

/* u01_open_open_64_53 (u01.h:129-131): (1 + (i >> 11)) * (2^53-1) * 2^-106 */
double orc_u01_open_open_64_53(uint64_t i)
{
    const double scale = 9007199254740991.0 * (1.0 / (4294967296.0 * 2097152.0)) *
                         (1.0 / (4294967296.0 * 2097152.0));
    return (double)(1 + (i >> 11)) * scale;
}

/* The reference stores the per-orbit RNG counter as a double and assigns it
 * to a 64-bit unsigned counter word (stochastic_solver_caller.cl:49).  For
 * non-negative values below 2^64 that is truncation; for negative values the
 * conversion is undefined in C, and this restatement (like the MI355X
 * integrators) takes the two's-complement of the truncated integer. */
uint64_t orc_counter_word(double c)
{
    if (c >= 0) {
        return (uint64_t)c;
    }
    return (uint64_t)(int64_t)c;
}

/* ------------------------------------------------------------------------ */
/* Solver steps                                                             */
/* ------------------------------------------------------------------------ */

/* euler.cl:10-18 */
static void orc_euler(double dt, double t, double* y, double* yout, double* p)
{
    sode_system(t, y, yout, p);
    for (int ieq = 0; ieq < _numeq_; ieq++) {
        yout[ieq] = y[ieq] + yout[ieq] * dt;
    }
}

/* rk4.cl:17-64.  k1 doubles as the stage-derivative buffer; yout collects
 * k1 + 2 k2 + 2 k3 and the update is y + (acc + k4) * dt / 6.00. */
static void orc_rk4(double dt, double t, double* y, double* yout, double* p)
{
    double k1[_numeq_];
    double x2[_numeq_];
    sode_system(t, y, yout, p);
    for (int i = 0; i < _numeq_; ++i) k1[i] = yout[i];
    for (int i = 0; i < _numeq_; ++i) x2[i] = y[i] + (dt * 0.50) * k1[i];
    sode_system(t + 0.50 * dt, x2, k1, p);
    for (int i = 0; i < _numeq_; ++i) yout[i] = yout[i] + 2.00 * k1[i];
    for (int i = 0; i < _numeq_; ++i) x2[i] = y[i] + (dt * 0.50) * k1[i];
    sode_system(t + 0.50 * dt, x2, k1, p);
    for (int i = 0; i < _numeq_; ++i) yout[i] = yout[i] + 2.00 * k1[i];
    for (int i = 0; i < _numeq_; ++i) x2[i] = y[i] + dt * k1[i];
    sode_system(t + dt, x2, k1, p);
    for (int i = 0; i < _numeq_; ++i) yout[i] = y[i] + (yout[i] + k1[i]) * dt / 6.00;
}

/* The reference's "Thomas" solve (ie.cl:10-48, im.cl:3-41).  It is a
 * textbook Thomas sweep over the tridiagonal matrix whose sub-diagonal is
 * a[i] = J[i][i+1] (zero for i >= N-2), diagonal b[i] = J[i][i] and
 * super-diagonal c[i] = J[i+1][i]; v is overwritten. */
static void orc_thomas(int n, const double* lower, const double* diag, const double* upper,
                       double* v, double* x)
{
    double a[n];
    double b[n];
    double c[n];
    for (int i = 0; i < n; i++) a[i] = 0;
    for (int i = 0; i < n; i++) c[i] = 0;
#ifdef ORC_PLANT_TEXTBOOK_THOMAS
    /* planted mistake for the oracle's own tests: the textbook assignment
     * (sub-diagonal a[i] = J[i][i-1], super-diagonal c[i] = J[i][i+1]) */
    for (int i = 1; i < n; i++) a[i] = lower[i - 1];
    for (int i = 0; i < n - 1; i++) c[i] = upper[i];
#else
    for (int i = 0; i < n - 2; i++) a[i] = upper[i];
    for (int i = 0; i < n - 1; i++) c[i] = lower[i];
#endif
    for (int i = 0; i < n; i++) b[i] = diag[i];
    for (int i = 1; i < n; i++) {
        double m = a[i] / b[i - 1];
        b[i] = b[i] - m * c[i - 1];
        v[i] = v[i] - m * v[i - 1];
    }
    x[n - 1] = v[n - 1] / b[n - 1];
    for (int i = n - 2; i >= 0; i--) {
        x[i] = (v[i] - c[i] * x[i + 1]) / b[i];
    }
}

/* Builds I + scale * J from the dense plugin Jacobian (ie.cl:83-92,
 * im.cl:84-93: jac[i][j] = jac[i][j] * scale_sign * dt, then +1 on the
 * diagonal) and returns its three diagonals. */
static void orc_newton_matrix(double* g, double* p, double sign, double dt,
                              double* lower, double* diag, double* upper)
{
    int n = _numeq_;
    if (orc_use_band && calc_jacobian_band) {
        calc_jacobian_band(g, lower, diag, upper, p);
        for (int i = 0; i < n; i++) {
            lower[i] = (i + 1 < n) ? lower[i] * sign * dt : 0;
            upper[i] = (i + 1 < n) ? upper[i] * sign * dt : 0;
            diag[i] = diag[i] * sign * dt;
            diag[i] = diag[i] + 1;
        }
        return;
    }
    double (*jac)[n] = malloc(sizeof(double) * (size_t)n * (size_t)n);
    calc_jacobian(g, jac, p);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            jac[i][j] = jac[i][j] * sign * dt;
        }
        jac[i][i] = jac[i][i] + 1;
    }
    for (int i = 0; i < n; i++) {
        diag[i] = jac[i][i];
        upper[i] = (i + 1 < n) ? jac[i][i + 1] : 0;
        lower[i] = (i + 1 < n) ? jac[i + 1][i] : 0;
    }
    free(jac);
}

/* ie.cl:50-117.  Newton on g - y - dt f(t, g) = 0 from g = y, stopping when
 * the L1 norm of the residual drops below _epsilon1_.  Returns the number
 * of Newton updates taken. */
static int orc_ie(double dt, double t, double* y, double* yout, double* p)
{
    int n = _numeq_;
    double guess[n], fval[n], res[n], corr[n], lo[n], di[n], up[n];
    int updates = 0;
    for (int i = 0; i < n; i++) guess[i] = y[i];
    for (int it = 0; it < orc_newton_maxit; it++) {
        sode_system(t, guess, fval, p);
        for (int i = 0; i < n; i++) res[i] = guess[i] - y[i] - fval[i] * dt;
        double err1 = 0;
        for (int i = 0; i < n; i++) err1 = err1 + fabs(res[i]);
        if (err1 < _epsilon1_) break;
        orc_newton_matrix(guess, p, -1.00, dt, lo, di, up);
        orc_thomas(n, lo, di, up, res, corr);
        for (int i = 0; i < n; i++) guess[i] = guess[i] - corr[i];
        updates++;
    }
    for (int i = 0; i < n; i++) yout[i] = guess[i];
    return updates;
}

/* im.cl:43-118.  Newton on g - y - dt f(t, (y+g)/2) = 0 from g = y, L2 stop
 * test, Jacobian at the midpoint scaled by -0.50 dt. */
static int orc_im(double dt, double t, double* y, double* yout, double* p)
{
    int n = _numeq_;
    double guess[n], mid[n], fval[n], res[n], corr[n], lo[n], di[n], up[n];
    int updates = 0;
    for (int i = 0; i < n; i++) guess[i] = y[i];
    for (int it = 0; it < orc_newton_maxit; it++) {
        for (int i = 0; i < n; i++) mid[i] = (y[i] + guess[i]) * 0.5;
        sode_system(t, mid, fval, p);
        for (int i = 0; i < n; i++) res[i] = guess[i] - y[i] - fval[i] * dt;
        double err1 = 0;
        for (int i = 0; i < n; i++) err1 = err1 + res[i] * res[i];
        if (sqrt(err1) < _epsilon1_) break;
        orc_newton_matrix(mid, p, -0.50, dt, lo, di, up);
        orc_thomas(n, lo, di, up, res, corr);
        for (int i = 0; i < n; i++) guess[i] = guess[i] - corr[i];
        updates++;
    }
    for (int i = 0; i < n; i++) yout[i] = guess[i];
    return updates;
}

/* stochastic_euler.cl:10-22 */
static void orc_em(double t, double* y, double* yout, double* p, double* noise)
{
    double stoch[_numeq_];
    sode_system(t, y, yout, p);
    sode_system_stoch(t, y, stoch, p, noise);
    for (int i = 0; i < _numeq_; i++) {
        yout[i] = y[i] + yout[i] * _m_dt_ + stoch[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Macro contract and host segment loop                                     */
/* ------------------------------------------------------------------------ */

/* double2string (src/sodeclmgr.cpp:356-367) prints with ostream defaults,
 * i.e. 6 significant digits (%g); the kernel sees that literal as _m_dt_. */
double orc_kernel_dt(double dt)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%g", dt);
    return strtod(buf, NULL);
}

/* Launch count of run_sode_solver (src/sodeclmgr.cpp:951): the number of
 * integers j >= 0 with j < tspan / (dt * ksteps), evaluated in double. */
long orc_num_launches(double tspan, double dt, int ksteps)
{
    double x = tspan / (dt * ksteps);
    if (!(x > 0)) return 0;
    if (x > 2147483647.0) return -1;
    long n = (long)ceil(x);
    while (n > 0 && !((double)(n - 1) < x)) n--;
    while ((double)n < x) n++;
    return n;
}

typedef struct {
    int solver;
    int neq, npar, nnoi, orbits, ksteps;
    double dt, tspan;
    int newton_maxit;
    int nthreads;
    long nlaunch;      /* >= 0: number of ksteps segments, overriding tspan */
    int use_band;      /* build Newton matrices from calc_jacobian_band */
    double newton_tol; /* > 0: Newton stop threshold in place of _epsilon1_ = 1e-07 */
} orc_config;

/* Integrates one orbit for one segment of ksteps steps, exactly as one
 * work-item of solver_caller.cl:8-51 / stochastic_solver_caller.cl:8-81.
 * counter is that orbit's double RNG counter (SDE only).  newton_updates
 * accumulates the Newton updates taken (implicit solvers). */
static void orc_segment(const orc_config* c, int orbit, double* t, double* y, double* p,
                        double* counter, long* newton_updates)
{
    int n = c->neq;
    double det[n];
    if (c->solver == ORC_SEULER) {
        double noise[c->nnoi > 0 ? c->nnoi : 1];
        double dtsqrt = sqrt(_m_dt_);
        uint64_t key[2] = {0, 0xf00dcafeULL};   /* rk.v[0] = 0 overwrites the id (:31,:38) */
        int calls = (int)ceilf((float)((float)(_numnoi_) * 0.5));
        for (int it = 0; it < _numsteps_; it++) {
            int nc = 0;
            double u0 = 0, u1 = 0;
            for (int in = 0; in < calls; in++) {
                uint64_t ctr[2] = {(uint64_t)orbit, orc_counter_word(*counter)};
                uint64_t rr[2];
                orc_threefry2x64_20(ctr, key, rr);
                *counter = *counter + 1;
                if (nc < _numnoi_) {
                    u0 = orc_u01_open_open_64_53(rr[0]);
                    u1 = orc_u01_open_open_64_53(rr[1]);
                    noise[nc] = (sqrt(-2 * log(u0)) * cos(2 * M_PI * u1)) * dtsqrt;
                    nc = nc + 1;
                }
                if (nc < _numnoi_) {
                    noise[nc] = (sqrt(-2 * log(u0)) * sin(2 * M_PI * u1)) * dtsqrt;
                    nc = nc + 1;
                }
            }
            orc_em(*t, y, det, p, noise);
            *t = *t + _m_dt_;
            for (int i = 0; i < n; i++) y[i] = det[i];
        }
        return;
    }
    for (int it = 0; it < _numsteps_; ++it) {
        switch (c->solver) {
        case ORC_EULER: orc_euler(_m_dt_, *t, y, det, p); break;
        case ORC_RK4: orc_rk4(_m_dt_, *t, y, det, p); break;
        case ORC_IE: *newton_updates += orc_ie(_m_dt_, *t, y, det, p); break;
        case ORC_IM: *newton_updates += orc_im(_m_dt_, *t, y, det, p); break;
        default: break;
        }
        *t = *t + _m_dt_;
        for (int i = 0; i < n; i++) y[i] = det[i];
    }
}

static void orc_set_globals(const orc_config* c)
{
    orc_numeq = c->neq;
    orc_numpar = c->npar;
    orc_numnoi = c->nnoi;
    orc_listsize = c->orbits;
    orc_numsteps = c->ksteps;
    orc_kdt = orc_kernel_dt(c->dt);
    orc_newton_maxit = c->newton_maxit > 0 ? c->newton_maxit : 1000;
    orc_use_band = c->use_band;
    orc_epsilon1 = c->newton_tol > 0 ? c->newton_tol : 1e-07;
}

/*
 * Full run as sodeclmgr::run_sode_solver (src/sodeclmgr.cpp:921-1035):
 * snapshot j is the state before launch j, plus one final snapshot; output
 * is flat [snapshot][equation][orbit] (:979-995).
 *   t0       [orbits]           in/out
 *   y        [orbits][neq]      in/out (AoS, as the reference's y0 buffer)
 *   params   [orbits][npar]
 *   counters [orbits] or NULL   in/out (SDE only)
 *   out      [nsnap][neq][orbits] or NULL
 *   newton_updates [orbits] or NULL: Newton updates taken per orbit
 * orbit_begin/orbit_count select a sub-range of orbits (the orbit id that
 * keys the RNG stays the global one); out then holds only that range.
 * Returns the number of launches, or -1 on bad input.
 */
long orc_run(const orc_config* c, int orbit_begin, int orbit_count, double* t0, double* y,
             const double* params, double* counters, double* out, long* newton_updates)
{
    if (c->neq <= 0 || c->npar < 0 || c->orbits <= 0 || c->ksteps <= 0) return -1;
    if (c->solver < 0 || c->solver > 4) return -1;
    if (c->solver == ORC_SEULER && (c->nnoi <= 0 || counters == NULL || !sode_system_stoch)) return -1;
    if ((c->solver == ORC_IE || c->solver == ORC_IM) && !calc_jacobian) return -1;
    orc_set_globals(c);
    long launches = c->nlaunch >= 0 ? c->nlaunch : orc_num_launches(c->tspan, c->dt, c->ksteps);
    if (launches < 0) return -1;
    int n = c->neq;
    int ob = orbit_begin, oc = orbit_count;
    long slab = (long)n * oc;
#ifdef _OPENMP
    int nth = c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nth)
#endif
    for (int o = 0; o < oc; o++) {
        int orbit = ob + o;
        double yy[n];
        double pp[c->npar > 0 ? c->npar : 1];
        for (int m = 0; m < n; m++) yy[m] = y[(long)orbit * n + m];
        for (int m = 0; m < c->npar; m++) pp[m] = params[(long)orbit * c->npar + m];
        double t = t0[orbit];
        double cnt = counters ? counters[orbit] : 0;
        long nu = 0;
        for (long j = 0; j <= launches; j++) {
            if (out) {
                for (int m = 0; m < n; m++) out[j * slab + (long)m * oc + o] = yy[m];
            }
            if (j == launches) break;
            orc_segment(c, orbit, &t, yy, pp, &cnt, &nu);
        }
        for (int m = 0; m < n; m++) y[(long)orbit * n + m] = yy[m];
        t0[orbit] = t;
        if (counters) counters[orbit] = cnt;
        if (newton_updates) newton_updates[orbit] = nu;
    }
    return launches;
}

/* One right-hand-side evaluation through the plugin (tests). */
void orc_rhs(int neq, int npar, int nnoi, double t, double* y, double* yout, double* p)
{
    orc_numeq = neq;
    orc_numpar = npar;
    orc_numnoi = nnoi;
    sode_system(t, y, yout, p);
}

/* One Box-Muller noise pair for a given counter word (tests). */
void orc_noise_pair(uint64_t orbit, uint64_t counter_word, double dt, double out[2])
{
    uint64_t key[2] = {0, 0xf00dcafeULL};
    uint64_t ctr[2] = {orbit, counter_word};
    uint64_t rr[2];
    orc_threefry2x64_20(ctr, key, rr);
    double dts = sqrt(orc_kernel_dt(dt));
    double u0 = orc_u01_open_open_64_53(rr[0]);
    double u1 = orc_u01_open_open_64_53(rr[1]);
    out[0] = (sqrt(-2 * log(u0)) * cos(2 * M_PI * u1)) * dts;
    out[1] = (sqrt(-2 * log(u0)) * sin(2 * M_PI * u1)) * dts;
}

/* ------------------------------------------------------------------------ */
/* Forward sensitivity (tangent-linear) of the explicit solvers: the FP64    */
/* reference for the MI355X discrete adjoint.  No reference counterpart: the */
/* reference computes no derivatives.                                        */
/* ------------------------------------------------------------------------ */

/* dk = J(x) dx + [0 | df/dp(x)], all (neq x ncol) with ncol = neq + npar */
static void orc_tangent_rhs(double t, double* x, double* p, const double* dx, double* dk, int ncol)
{
    int n = _numeq_, np = _numpar_;
    double (*jac)[n] = malloc(sizeof(double) * (size_t)n * (size_t)n);
    double (*jp)[np > 0 ? np : 1] = malloc(sizeof(double) * (size_t)n * (size_t)(np > 0 ? np : 1));
    (void)t;
    /* the derivative of the right-hand side: the exact Jacobian where the
     * plugin has one (the oculomotor's calc_jacobian keeps the reference's
     * Newton-matrix sign in J[4][5], which is not df/dy) */
    if (calc_jacobian_exact) calc_jacobian_exact(x, jac, p);
    else calc_jacobian(x, jac, p);
    if (np > 0) calc_jacobian_params(x, jp, p);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < ncol; k++) {
            double acc = 0;
            for (int m = 0; m < n; m++) acc += jac[i][m] * dx[(size_t)m * ncol + k];
            if (k >= n) acc += jp[i][k - n];
            dk[(size_t)i * ncol + k] = acc;
        }
    }
    free(jac);
    free(jp);
}

/* The plugin's Jacobians at (y, p) for the checks of tests: jac [neq][neq]
 * from calc_jacobian_exact when exact and the plugin has it, else
 * calc_jacobian; jp [neq][npar] from calc_jacobian_params.  Returns -1 when
 * an entry point is missing. */
long orc_jacobians(int neq, int npar, double* y, double* p, double* jac, double* jp, int exact)
{
    orc_numeq = neq;
    orc_numpar = npar;
    if (!calc_jacobian || (npar > 0 && !calc_jacobian_params)) return -1;
    if (exact && calc_jacobian_exact) calc_jacobian_exact(y, (double (*)[orc_numeq])jac, p);
    else calc_jacobian(y, (double (*)[orc_numeq])jac, p);
    if (npar > 0) calc_jacobian_params(y, (double (*)[orc_numpar])jp, p);
    return 0;
}

/* Solves A X = B in place (A n x n row-major, B n x ncol row-major) by
 * Gaussian elimination with partial pivoting; X replaces B. */
static void orc_dense_solve(int n, double* A, double* B, int ncol)
{
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[(size_t)i * n + k]) > fabs(A[(size_t)piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; j++) {
                double tmp = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)piv * n + j];
                A[(size_t)piv * n + j] = tmp;
            }
            for (int j = 0; j < ncol; j++) {
                double tmp = B[(size_t)k * ncol + j];
                B[(size_t)k * ncol + j] = B[(size_t)piv * ncol + j];
                B[(size_t)piv * ncol + j] = tmp;
            }
        }
        for (int i = k + 1; i < n; i++) {
            double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            if (f == 0) continue;
            for (int j = k; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
            for (int j = 0; j < ncol; j++) B[(size_t)i * ncol + j] -= f * B[(size_t)k * ncol + j];
        }
    }
    for (int k = n - 1; k >= 0; k--)
        for (int j = 0; j < ncol; j++) {
            double a = B[(size_t)k * ncol + j];
            for (int m = k + 1; m < n; m++) a -= A[(size_t)k * n + m] * B[(size_t)m * ncol + j];
            B[(size_t)k * ncol + j] = a / A[(size_t)k * n + k];
        }
}

/* Tangent of one implicit step (IE: c = 1, x = g; IM: c = 1/2,
 * x = (y + g)/2) by the implicit function theorem at the converged Newton
 * solution g:
 *   IE: (I - dt J(g)) S' = S + dt [0 | f_p(g)]
 *   IM: (I - dt/2 J(m)) S' = (I + dt/2 J(m)) S + dt [0 | f_p(m)]
 * with J the plugin's exact df/dy (calc_jacobian_exact, else calc_jacobian).
 * S is n x ncol (ncol = n + npar), updated in place. */
static void orc_implicit_tangent(int solver, double dt, const double* y, const double* g, double* p, double* S,
                                 int ncol)
{
    int n = _numeq_, np = _numpar_;
    double c = solver == ORC_IE ? 1.0 : 0.5;
    double x[n];
    for (int i = 0; i < n; i++) x[i] = solver == ORC_IE ? g[i] : (y[i] + g[i]) * 0.5;
    double (*jac)[n] = malloc(sizeof(double) * (size_t)n * (size_t)n);
    double (*jp)[np > 0 ? np : 1] = malloc(sizeof(double) * (size_t)n * (size_t)(np > 0 ? np : 1));
    if (calc_jacobian_exact) calc_jacobian_exact(x, jac, p);
    else calc_jacobian(x, jac, p);
    if (np > 0) calc_jacobian_params(x, jp, p);
    double* A = malloc(sizeof(double) * (size_t)n * (size_t)n);
    double* B = malloc(sizeof(double) * (size_t)n * (size_t)ncol);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = (i == j ? 1.0 : 0.0) - c * dt * jac[i][j];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < ncol; k++) {
            double b = S[(size_t)i * ncol + k];
            if (solver == ORC_IM) {
                double a = 0;
                for (int m = 0; m < n; m++) a += jac[i][m] * S[(size_t)m * ncol + k];
                b += 0.5 * dt * a;
            }
            if (k >= n) b += dt * jp[i][k - n];
            B[(size_t)i * ncol + k] = b;
        }
    orc_dense_solve(n, A, B, ncol);
    memcpy(S, B, sizeof(double) * (size_t)n * (size_t)ncol);
    free(A);
    free(B);
    free(jac);
    free(jp);
}

/*
 * Gradient of L = sum_snap <cot[snap], y(snap)> with respect to y0 and the
 * parameters of each orbit in [orbit_begin, orbit_begin + orbit_count), by
 * propagating S = d y / d(y0, p) through the same discrete steps as the
 * forward integration: the exact tangent of the Euler or RK4 step, or for
 * implicit Euler / midpoint the tangent of each step's implicit equation at
 * the converged Newton solution (orc_implicit_tangent).  cot is
 * [nsnap][neq][orbits] over all orbits; grad_y0 [count][neq], grad_p
 * [count][npar].  Returns 0, or -1 when the plugin lacks calc_jacobian /
 * calc_jacobian_params or the solver is stochastic.
 */
long orc_sensitivity_gradient(const orc_config* c, int orbit_begin, int orbit_count, const double* y0,
                              const double* params, const double* cot, double* grad_y0, double* grad_p)
{
    if (!calc_jacobian || (c->npar > 0 && !calc_jacobian_params)) return -1;
    if (c->solver == ORC_SEULER || c->solver < 0 || c->solver > 4) return -1;
    orc_set_globals(c);
    long launches = c->nlaunch >= 0 ? c->nlaunch : orc_num_launches(c->tspan, c->dt, c->ksteps);
    int n = c->neq, np = c->npar, ncol = n + np;
    double dt = _m_dt_;
#ifdef _OPENMP
    int nth = c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
#endif
    for (int o = 0; o < orbit_count; o++) {
        int orbit = orbit_begin + o;
        double* S = calloc((size_t)n * ncol, sizeof(double));
        double* dk = calloc((size_t)n * ncol, sizeof(double));
        double* acc = calloc((size_t)n * ncol, sizeof(double));
        double* dxs = calloc((size_t)n * ncol, sizeof(double));
        double* g = calloc((size_t)ncol, sizeof(double));
        double y[n], pp[np > 0 ? np : 1], f[n], k[n], xs[n], yo[n];
        for (int i = 0; i < n; i++) {
            y[i] = y0[(long)orbit * n + i];
            S[(size_t)i * ncol + i] = 1.0;
        }
        for (int m = 0; m < np; m++) pp[m] = params[(long)orbit * np + m];
        double t = 0;
        long total = launches * c->ksteps;
        for (long s = 0; s <= total; s++) {
            if (s % c->ksteps == 0) {
                const double* cs = cot + (s / c->ksteps) * (long)n * c->orbits;
                for (int col = 0; col < ncol; col++) {
                    double a = 0;
                    for (int i = 0; i < n; i++) a += cs[(long)i * c->orbits + orbit] * S[(size_t)i * ncol + col];
                    g[col] += a;
                }
            }
            if (s == total) break;
            if (c->solver == ORC_IE || c->solver == ORC_IM) {
                double g[n];
                if (c->solver == ORC_IE) orc_ie(dt, t, y, g, pp);
                else orc_im(dt, t, y, g, pp);
                orc_implicit_tangent(c->solver, dt, y, g, pp, S, ncol);
                for (int i = 0; i < n; i++) y[i] = g[i];
            } else if (c->solver == ORC_EULER) {
                orc_tangent_rhs(t, y, pp, S, dk, ncol);
                sode_system(t, y, f, pp);
                for (size_t q = 0; q < (size_t)n * ncol; q++) S[q] = S[q] + dk[q] * dt;
                for (int i = 0; i < n; i++) y[i] = y[i] + f[i] * dt;
            } else {
                /* stage 1 */
                orc_tangent_rhs(t, y, pp, S, dk, ncol);
                sode_system(t, y, f, pp);
                for (size_t q = 0; q < (size_t)n * ncol; q++) { acc[q] = dk[q]; dxs[q] = S[q] + (dt * 0.5) * dk[q]; }
                for (int i = 0; i < n; i++) { yo[i] = f[i]; xs[i] = y[i] + (dt * 0.50) * f[i]; }
                /* stage 2 */
                orc_tangent_rhs(t + 0.50 * dt, xs, pp, dxs, dk, ncol);
                sode_system(t + 0.50 * dt, xs, k, pp);
                for (size_t q = 0; q < (size_t)n * ncol; q++) { acc[q] += 2 * dk[q]; dxs[q] = S[q] + (dt * 0.5) * dk[q]; }
                for (int i = 0; i < n; i++) { yo[i] = yo[i] + 2.00 * k[i]; xs[i] = y[i] + (dt * 0.50) * k[i]; }
                /* stage 3 */
                orc_tangent_rhs(t + 0.50 * dt, xs, pp, dxs, dk, ncol);
                sode_system(t + 0.50 * dt, xs, k, pp);
                for (size_t q = 0; q < (size_t)n * ncol; q++) { acc[q] += 2 * dk[q]; dxs[q] = S[q] + dt * dk[q]; }
                for (int i = 0; i < n; i++) { yo[i] = yo[i] + 2.00 * k[i]; xs[i] = y[i] + dt * k[i]; }
                /* stage 4 */
                orc_tangent_rhs(t + dt, xs, pp, dxs, dk, ncol);
                sode_system(t + dt, xs, k, pp);
                for (size_t q = 0; q < (size_t)n * ncol; q++) S[q] = S[q] + (acc[q] + dk[q]) * dt / 6.0;
                for (int i = 0; i < n; i++) y[i] = y[i] + (yo[i] + k[i]) * dt / 6.00;
            }
            t = t + dt;
        }
        for (int i = 0; i < n; i++) grad_y0[(long)o * n + i] = g[i];
        for (int m = 0; m < np; m++) grad_p[(long)o * np + m] = g[n + m];
        free(S); free(dk); free(acc); free(dxs); free(g);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Check of the division identity the MI355X code and the bundled plugins   */
/* use for divisions by constants and by orbit parameters: with rd = RN(1/d)*/
/* q0 = RN(a rd), r = a - q0 d (exact, one fma), q1 = RN(q0 + r rd) equals   */
/* RN(a / d) (Markstein); zeros, inf, NaN pass as in the kernels.  Counts   */
/* mismatches over n random normal dividends plus special values, and those */
/* of the select-free form against it (subnormal and overflowing quotients, */
/* infinities and NaN included) within that form's contract.                */
/* ------------------------------------------------------------------------ */
static double orc_markstein(double a, double d, double rd)
{
    const double q0 = a * rd;
    const double r = fma(-q0, d, a);
    const double q1 = fma(r, rd, q0);
    return (isfinite(q0) && q0 != 0.0) ? q1 : q0;
}

/* The select-free form the kernels and the oculomotor plugin use: residual */
/* q0 d - a clamped below at -DBL_MAX.  Its contract: d >= 1, or a dividend  */
/* whose sign bit is set.                                                    */
static double orc_markstein_clamped(double a, double d, double rd)
{
    const double q0 = a * rd;
    const double r = fmax(fma(q0, d, -a), -DBL_MAX);
    return fma(-r, rd, q0);
}

static int orc_clamped_agrees(double a, double d, double rd)
{
    if (!(d >= 1.0 || signbit(a))) return 1;
    double g = orc_markstein(a, d, rd), c = orc_markstein_clamped(a, d, rd);
    if (isnan(g) || isnan(c)) return isnan(g) && isnan(c);
    return memcmp(&g, &c, 8) == 0;
}

long orc_markstein_mismatches(double d, long n, uint64_t seed)
{
    uint64_t s0 = seed ^ 0x9E3779B97F4A7C15ULL, s1 = seed * 0xBF58476D1CE4E5B9ULL + 1;
    const double rd = 1.0 / d;
    long bad = 0;
    const double special[] = {0.0, -0.0, INFINITY, -INFINITY, 1.0, -1.0, d, -d, 3.0 * d, 1e300, -1e300, 1e-300};
    for (size_t i = 0; i < sizeof special / sizeof special[0]; i++) {
        double q = special[i] / d, m = orc_markstein(special[i], d, rd);
        if (memcmp(&q, &m, 8) != 0) bad++;
        if (!orc_clamped_agrees(special[i], d, rd) || !orc_clamped_agrees(-special[i], d, rd)) bad++;
    }
    const double edge[] = {DBL_MAX, 1e-310, 5e-324, 2.2250738585072014e-308, 1e-300 * d, DBL_MAX * (d < 1 ? d : 1)};
    for (size_t i = 0; i < sizeof edge / sizeof edge[0]; i++)
        if (!orc_clamped_agrees(edge[i], d, rd) || !orc_clamped_agrees(-edge[i], d, rd)) bad++;
    if (!orc_clamped_agrees(NAN, d, rd) || !orc_clamped_agrees(-NAN, d, rd)) bad++;
    double nq = NAN / d, nm = orc_markstein(NAN, d, rd);
    if (!(isnan(nq) && isnan(nm))) bad++;
    for (long i = 0; i < n; i++) {
        uint64_t x = s0, y = s1;   /* xorshift128+ */
        s0 = y;
        x ^= x << 23;
        s1 = x ^ y ^ (x >> 17) ^ (y >> 26);
        uint64_t m = s1 + y;
        uint64_t e = ((m >> 52) % 1600) + 200;   /* exponents well inside the normal range */
        uint64_t bits = (m & 0x800FFFFFFFFFFFFFULL) | (e << 52);
        double a;
        memcpy(&a, &bits, 8);
        double q = a / d, mm = orc_markstein(a, d, rd);
        if (memcmp(&q, &mm, 8) != 0) bad++;
        if (!orc_clamped_agrees(a, d, rd) || !orc_clamped_agrees(-a, d, rd)) bad++;
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* On-device sampling (sodecl_sample_device): entry (r, c) of out[rows][cols] */
/* from Threefry-2x64-20 with counter {r, c}, key {seed, 0x5EEDD15C0FFEE5ED}; */
/* kind 0 constant a, 1 uniform a + (b-a) u0 (one fma), 2 log-uniform        */
/* exp(log a + (log b - log a) u0), 3 normal a + b sqrt(-2 log u0) cos(2 pi  */
/* u1).  Restated from the definition the library documents.                */
/* ------------------------------------------------------------------------ */
void orc_sample(const int* kind, const double* a, const double* b, int rows, int cols, uint64_t seed, double* out)
{
    const uint64_t key[2] = {seed, 0x5EEDD15C0FFEE5EDULL};
    for (long r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            double v = a[c];
            if (kind[c] != 0) {
                uint64_t ctr[2] = {(uint64_t)r, (uint64_t)c}, w[2];
                orc_threefry2x64_20(ctr, key, w);
                double u0 = orc_u01_open_open_64_53(w[0]);
                if (kind[c] == 1) {
                    v = fma(b[c] - a[c], u0, a[c]);
                } else if (kind[c] == 2) {
                    double la = log(a[c]);
                    v = exp(fma(log(b[c]) - la, u0, la));
                } else {
                    double u1 = orc_u01_open_open_64_53(w[1]);
                    v = fma(b[c], sqrt(-2.0 * log(u0)) * cos(2.0 * M_PI * u1), a[c]);
                }
            }
            out[r * cols + c] = v;
        }
}
