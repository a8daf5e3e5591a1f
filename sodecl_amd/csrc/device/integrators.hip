// SODECL MI355X integrators: device code compiled at run time by hipRTC for
// gfx950, after the macro prelude and the user's system source.
//
// Macros provided by the host (sodecl_amd/csrc/jit.cpp):
//   _numeq_ _numpar_ _numnoi_ _listsize_ _numsteps_ _m_dt_ [_epsilon1_]
//       the reference's macro contract (src/sodeclmgr.cpp:390-430)
//   SODECL_SOLVER          0 Euler-Maruyama, 1 Euler, 2 RK4, 3 implicit Euler,
//                          4 implicit midpoint (include/solver_Type.hpp:14-20)
//   SODECL_BLOCK           threads per block
//   SODECL_GROUP           lanes per orbit in the group schedule
//   SODECL_NEWTON_MAXIT    Newton iteration cap (ie.cl:69 uses 1000)
//   SODECL_STRICT          1: strict arithmetic (fp_contract = 1): no
//                          contraction and correctly rounded divisions in
//                          the solver steps; 0: contraction, and divisions
//                          by constants as products with the reciprocal
//   SODECL_HAS_*           optional plugin entry points found in the source
//
// Schedules:
//   sodecl_lane   one orbit per lane; state, parameters and stage vectors
//                 live in registers (small systems such as the oculomotor)
//   sodecl_group  SODECL_GROUP lanes per orbit; lane l owns components
//                 l, l+G, ...; stage vectors are staged in LDS and every lane
//                 evaluates its components through the plugin's per-component
//                 entry point (coupled systems such as Kuramoto)
//   sodecl_team   SODECL_GROUP lanes per orbit, each with the whole state;
//                 the plugin's team entry point shares the costly terms
// All run many ksteps segments per launch and write the snapshot that
// precedes each segment straight into the trajectory buffer, laid out
// [snapshot][equation][orbit] as the reference's output (sodeclmgr.cpp:979-995).

#ifndef SODECL_INLINE
#define SODECL_INLINE __device__ __attribute__((always_inline)) inline
#endif

#if _numeq_ <= 32
#define SODECL_UNROLL _Pragma("unroll")
#else
#define SODECL_UNROLL
#endif

#define SODECL_NPAR_ALLOC (_numpar_ > 0 ? _numpar_ : 1)
#define SODECL_NNOI_ALLOC (_numnoi_ > 0 ? _numnoi_ : 1)
// Threefry calls per step: one per pair of noise values
#define SODECL_CALLS ((_numnoi_ + 1) / 2)
#define SODECL_IS_SDE (SODECL_SOLVER == 0)
#define SODECL_IS_IMPLICIT (SODECL_SOLVER == 3 || SODECL_SOLVER == 4)


// a / d for a divisor d >= 1 whose correctly rounded reciprocal rd = 1/d is
// known (a constant, or hoisted out of the step loop): q0 = a*rd is within
// one ulp, the residual q0*d - a is exact in one fused multiply-add, and one
// more corrects q0 to the correctly rounded quotient (Markstein's theorem),
// the same bits as IEEE a / d outside the subnormal range.  Special values
// need no select: with the residual's sign taken this way a zero quotient
// keeps its sign; an infinite a gives a NaN residual, which the clamp to
// -DBL_MAX turns into a finite one so that the correction leaves q0 (the
// infinity) in place; a NaN a stays NaN through q0; d >= 1 rules out an
// overflowing q0.  Four full-rate FP64 operations instead of the
// quarter-rate reciprocal and scale/fixup sequence of a general division.
SODECL_INLINE double sodecl_div(double a, double d, double rd)
{
    const double q0 = a * rd;
    const double r = fmax(fma(q0, d, -a), -0x1.fffffffffffffp+1023);
    return fma(-r, rd, q0);
}

// (a dt) / 6.00 of the RK4 update (rk4.cl:60): under strict arithmetic the
// correctly rounded quotient, otherwise the product with the rounded
// reciprocal (within an ulp of it; the same size of rounding change as
// contraction, and right for infinities, NaN and signed zeros).
#if SODECL_STRICT
#define SODECL_RK4_DIV6(a) sodecl_div((a), 6.00, 1.0 / 6.00)
#else
#define SODECL_RK4_DIV6(a) ((a) * (1.0 / 6.00))
#endif

// The reference keeps the per-orbit counter as a double and assigns it to the
// 64-bit counter word (stochastic_solver_caller.cl:49).  Non-negative values
// truncate; negative values (undefined in C) take the two's complement of the
// truncated integer, as the CPU restatement does.
SODECL_INLINE sodecl_u64 sodecl_counter_word(double c)
{
    return c >= 0.0 ? (sodecl_u64)c : (sodecl_u64)(long long)c;
}

// Key and counter of the reference's SDE caller: key {0, 0xf00dcafe} (the
// orbit id written into the key is overwritten with 0, :31 and :38), counter
// {orbit, counter_g[orbit]}, one counter increment per Threefry call.
#define SODECL_RNG_KEY0 0ull
#define SODECL_RNG_KEY1 0xf00dcafeull

// One Box-Muller pair scaled by sqrt(dt) (stochastic_solver_caller.cl:54-65),
// with the logarithm, sine and cosine specialised to the uniforms' ranges
// (gauss.hip).
SODECL_INLINE void sodecl_noise_pair(int orbit, double counter, double dtsqrt, double* z0, double* z1)
{
    sodecl_u64 w0, w1;
    sodecl_threefry2x64_20((sodecl_u64)orbit, sodecl_counter_word(counter), SODECL_RNG_KEY0, SODECL_RNG_KEY1,
                           &w0, &w1);
    const double u0 = sodecl_u01(w0);
    const double u1 = sodecl_u01(w1);
    const double r = sqrt(-2 * sodecl_log_positive(u0));
    double s, c;
    sodecl_sincos_reduced(2 * M_PI * u1, &s, &c);
    *z0 = (r * c) * dtsqrt;
    *z1 = (r * s) * dtsqrt;
}

// ---------------------------------------------------------------------------
// Implicit solvers: Newton iteration with the reference's tridiagonal solve.
// ---------------------------------------------------------------------------
#if SODECL_IS_IMPLICIT

// Three diagonals of I + scale*J at g (ie.cl:83-92, im.cl:84-93: every entry
// multiplied by scale then by dt, then 1 added to the diagonal).
SODECL_INLINE void sodecl_newton_band(double* g, double* p, double scale, double* lo, double* di, double* up)
{
#if SODECL_HAS_JACOBIAN_BAND
    calc_jacobian_band(g, lo, di, up, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) {
        lo[i] = lo[i] * scale * _m_dt_;
        up[i] = up[i] * scale * _m_dt_;
        di[i] = di[i] * scale * _m_dt_;
        di[i] = di[i] + 1;
    }
#else
    double jac[_numeq_][_numeq_];
    calc_jacobian(g, jac, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) {
        di[i] = jac[i][i] * scale * _m_dt_;
        di[i] = di[i] + 1;
        up[i] = (i + 1 < _numeq_) ? jac[i][i + 1] * scale * _m_dt_ : 0.0;
        lo[i] = (i + 1 < _numeq_) ? jac[i + 1][i] * scale * _m_dt_ : 0.0;
    }
#endif
}

// The reference's "Thomas" sweep (ie.cl:10-48): a Thomas solve of the
// tridiagonal matrix with sub-diagonal a[i] = J[i][i+1] (zero for i >= N-2),
// diagonal b[i] = J[i][i] and super-diagonal c[i] = J[i+1][i]; v is consumed.
SODECL_INLINE void sodecl_thomas(const double* lo, const double* di, const double* up, double* v, double* x)
{
    double b[_numeq_];
    b[0] = di[0];
    SODECL_UNROLL
    for (int i = 1; i < _numeq_; ++i) {
        const double a = (i < _numeq_ - 2) ? up[i] : 0.0;
        const double m = a / b[i - 1];
        b[i] = di[i] - m * lo[i - 1];
        v[i] = v[i] - m * v[i - 1];
    }
    x[_numeq_ - 1] = v[_numeq_ - 1] / b[_numeq_ - 1];
    SODECL_UNROLL
    for (int i = _numeq_ - 2; i >= 0; --i) {
        x[i] = (v[i] - lo[i] * x[i + 1]) / b[i];
    }
}

// One implicit step from y: implicit Euler (ie.cl:50-117, L1 stop test,
// f evaluated at (t, g)) or implicit midpoint (im.cl:43-118, L2 stop test,
// f and J evaluated at (y + g)/2).  Returns the Newton updates taken.
SODECL_INLINE int sodecl_implicit_step(double t, double* y, double* yout, double* p)
{
    double g[_numeq_], r[_numeq_], lo[_numeq_], di[_numeq_], up[_numeq_], x[_numeq_];
    int updates = 0;
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) g[i] = y[i];
    for (int it = 0; it < SODECL_NEWTON_MAXIT; ++it) {
#if SODECL_SOLVER == 3
        sode_system(t, g, r, p);
        double err = 0;
        SODECL_UNROLL
        for (int i = 0; i < _numeq_; ++i) {
            r[i] = g[i] - y[i] - r[i] * _m_dt_;
            err = err + fabs(r[i]);
        }
        if (err < _epsilon1_) break;
        sodecl_newton_band(g, p, -1.00, lo, di, up);
#else
        double mid[_numeq_];
        SODECL_UNROLL
        for (int i = 0; i < _numeq_; ++i) mid[i] = (y[i] + g[i]) * 0.5;
        sode_system(t, mid, r, p);
        double err = 0;
        SODECL_UNROLL
        for (int i = 0; i < _numeq_; ++i) {
            r[i] = g[i] - y[i] - r[i] * _m_dt_;
            err = err + r[i] * r[i];
        }
        if (sqrt(err) < _epsilon1_) break;
        sodecl_newton_band(mid, p, -0.50, lo, di, up);
#endif
        sodecl_thomas(lo, di, up, r, x);
        SODECL_UNROLL
        for (int i = 0; i < _numeq_; ++i) g[i] = g[i] - x[i];
        ++updates;
    }
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) yout[i] = g[i];
    return updates;
}
#endif  // SODECL_IS_IMPLICIT

// ---------------------------------------------------------------------------
// One step of the selected solver with the whole state per lane (lane and
// team schedules).  Under the team schedule the drift comes from the
// plugin's team entry point, which the lanes of a team evaluate together.
// ---------------------------------------------------------------------------
// Optional attributes of the integration kernels (e.g. occupancy hints)
#ifndef SODECL_KERNEL_ATTR
#define SODECL_KERNEL_ATTR
#endif

#if SODECL_SCHEDULE == 4
#define SODECL_RHS sode_system_team
#else
#define SODECL_RHS sode_system
#endif
#if SODECL_SOLVER == 1
// Euler (euler.cl:10-18)
SODECL_INLINE int sodecl_lane_step(double t, double* y, double* yout, double* p)
{
    SODECL_RHS(t, y, yout, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) yout[i] = y[i] + yout[i] * _m_dt_;
    return 0;
}
#elif SODECL_SOLVER == 2
// Classic RK4 (rk4.cl:17-64): yout accumulates k1 + 2 k2 + 2 k3, the update
// is y + (acc + k4) * dt / 6.00, stage times t, t+dt/2, t+dt/2, t+dt.
SODECL_INLINE int sodecl_lane_step(double t, double* y, double* yout, double* p)
{
    double k[_numeq_], x[_numeq_];
    SODECL_RHS(t, y, yout, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) x[i] = y[i] + (_m_dt_ * 0.50) * yout[i];
    SODECL_RHS(t + 0.50 * _m_dt_, x, k, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) {
        yout[i] = yout[i] + 2.00 * k[i];
        x[i] = y[i] + (_m_dt_ * 0.50) * k[i];
    }
    SODECL_RHS(t + 0.50 * _m_dt_, x, k, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) {
        yout[i] = yout[i] + 2.00 * k[i];
        x[i] = y[i] + _m_dt_ * k[i];
    }
    SODECL_RHS(t + _m_dt_, x, k, p);
    SODECL_UNROLL
    for (int i = 0; i < _numeq_; ++i) yout[i] = y[i] + SODECL_RK4_DIV6((yout[i] + k[i]) * _m_dt_);
    return 0;
}
#elif SODECL_IS_IMPLICIT
SODECL_INLINE int sodecl_lane_step(double t, double* y, double* yout, double* p)
{
    return sodecl_implicit_step(t, y, yout, p);
}
#endif

// ---------------------------------------------------------------------------
// Lane schedule kernel: one orbit per lane.
//   t_state[orbits], y_state[orbits][neq] (AoS, the reference's y0 buffer),
//   params[orbits][npar], counters[orbits] (SDE), all updated in place;
//   traj: snapshot buffer [snap][neq][orbits] or null; snapshot s0 + s is the
//   state before segment s, and s0 + nseg the final state when final_snap.
//   newton: per-orbit Newton update counts (implicit solvers) or null.
// ---------------------------------------------------------------------------
#if SODECL_SCHEDULE == 1
// Orbit i of a lane-schedule launch: one lane of sodecl_lane, or one
// iteration of the host share's orbit loop (SODECL_HOST_BUILD).
SODECL_INLINE void sodecl_lane_orbit(int i, double* __restrict__ t_state, double* __restrict__ y_state,
                                     const double* __restrict__ params, double* __restrict__ counters,
                                     double* __restrict__ traj, long long* __restrict__ newton, int orbits, int s0,
                                     int nseg, int final_snap, double* __restrict__ ckpt_t,
                                     double* __restrict__ ckpt_y, int ckpt_every, long long gstep0)
{
    double y[_numeq_], yn[_numeq_], p[SODECL_NPAR_ALLOC];
    SODECL_UNROLL
    for (int m = 0; m < _numeq_; ++m) y[m] = y_state[(size_t)i * _numeq_ + m];
    for (int m = 0; m < _numpar_; ++m) p[m] = params[(size_t)i * _numpar_ + m];
    double t = t_state[i];
    const size_t slab = (size_t)_numeq_ * (size_t)orbits;
#if SODECL_IS_SDE
    double cnt = counters[i];
    const double dtsqrt = sqrt(_m_dt_);
    const int calls = (int)ceilf((float)((float)(_numnoi_) * 0.5));
    double noise[SODECL_NNOI_ALLOC];
#endif
    long long nu = 0;
    // checkpoint k is the state before global step k * ckpt_every
    long long ck = ckpt_y ? (gstep0 + ckpt_every - 1) / ckpt_every : 0;   // next checkpoint index
    long long ck_step = ck * ckpt_every;
    long long g = gstep0;
    for (int s = 0; s < nseg; ++s) {
        if (traj) {
            double* o = traj + (size_t)(s0 + s) * slab + i;
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) o[(size_t)m * orbits] = y[m];
        }
        // The segment's steps in runs that end at the next checkpoint, so that
        // the step loop tests only its count (with one wavefront per SIMD, each
        // scalar instruction in it takes an issue cycle from the VALU).
        for (int it = 0; it < _numsteps_;) {
            int run = _numsteps_ - it;
            if (ckpt_y) {
                if (g == ck_step) {
                    const long long k = ck++;
                    ck_step += ckpt_every;
                    SODECL_UNROLL
                    for (int m = 0; m < _numeq_; ++m) ckpt_y[(size_t)k * slab + (size_t)m * orbits + i] = y[m];
                    ckpt_t[(size_t)k * orbits + i] = t;
                }
                if (ck_step - g < run) run = (int)(ck_step - g);   // g < ck_step here
            }
            it += run;
            g += run;
            for (; run > 0; --run) {
#if SODECL_IS_SDE
                // Noise for this step (stochastic_solver_caller.cl:44-66)
                int nc = 0;
                for (int in = 0; in < calls; ++in) {
                    double z0, z1;
                    sodecl_noise_pair(i, cnt, dtsqrt, &z0, &z1);
                    cnt = cnt + 1;
                    if (nc < _numnoi_) noise[nc++] = z0;
                    if (nc < _numnoi_) noise[nc++] = z1;
                }
                // Euler-Maruyama (stochastic_euler.cl:10-22)
                {
                    double st[_numeq_];
                    sode_system(t, y, yn, p);
                    sode_system_stoch(t, y, st, p, noise);
                    SODECL_UNROLL
                    for (int m = 0; m < _numeq_; ++m) yn[m] = y[m] + yn[m] * _m_dt_ + st[m];
                }
#else
                nu += sodecl_lane_step(t, y, yn, p);
#endif
                t = t + _m_dt_;
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m) y[m] = yn[m];
            }
        }
    }
    if (traj && final_snap) {
        double* o = traj + (size_t)(s0 + nseg) * slab + i;
        SODECL_UNROLL
        for (int m = 0; m < _numeq_; ++m) o[(size_t)m * orbits] = y[m];
    }
    SODECL_UNROLL
    for (int m = 0; m < _numeq_; ++m) y_state[(size_t)i * _numeq_ + m] = y[m];
    t_state[i] = t;
#if SODECL_IS_SDE
    counters[i] = cnt;
#endif
    if (newton) newton[i] += nu;
}

#if SODECL_HOST_BUILD
// The host share: orbits [first, last) of the same arrays (host memory, the
// full orbit count's layouts), on the calling thread and its OpenMP team.
extern "C" void sodecl_host_lane(double* t_state, double* y_state, const double* params, double* counters,
                                 double* traj, long long* newton, int orbits, int first, int last, int s0, int nseg,
                                 int final_snap, long long gstep0, int nthreads)
{
    // one orbit per work item: an orbit is a whole run of steps, and a share
    // of a few orbits per thread must still spread over every thread
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = first; i < last; ++i)
        sodecl_lane_orbit(i, t_state, y_state, params, counters, traj, newton, orbits, s0, nseg, final_snap, nullptr,
                          nullptr, 1, gstep0);
}
#else
extern "C" __global__ void __launch_bounds__(SODECL_BLOCK) SODECL_KERNEL_ATTR
sodecl_lane(double* __restrict__ t_state, double* __restrict__ y_state, const double* __restrict__ params,
            double* __restrict__ counters, double* __restrict__ traj, long long* __restrict__ newton,
            int orbits, int s0, int nseg, int final_snap, double* __restrict__ ckpt_t,
            double* __restrict__ ckpt_y, int ckpt_every, long long gstep0, double* __restrict__ stage_ws,
            long long ws_steps)
{
    SODECL_EXP_TABLE_INIT();
    const int i = blockIdx.x * SODECL_BLOCK + threadIdx.x;
    if (i >= orbits) return;
    (void)stage_ws;
    (void)ws_steps;
    sodecl_lane_orbit(i, t_state, y_state, params, counters, traj, newton, orbits, s0, nseg, final_snap, ckpt_t,
                      ckpt_y, ckpt_every, gstep0);
}
#endif
#endif  // SODECL_SCHEDULE == 1

// ---------------------------------------------------------------------------
// Team schedule kernel: SODECL_GROUP lanes per orbit, each holding the whole
// state and stepping it with the same arithmetic as the lane schedule; the
// plugin's sode_system_team shares the drift's costly terms among the team.
// This multiplies the wavefronts of a launch by G at the cost of the work the
// team duplicates, which pays only when the shared terms dominate the step
// and the orbits leave SIMDs idle (for the oculomotor model, whose two burst
// curves are about half of an RK4 step, it does not: 4.4e10 against 6.1e10
// orbit-steps/s for 65536 orbits).  Lane l of a team stores components m with
// m % G == l.  Arguments as sodecl_lane (explicit, non-stochastic solvers).
// ---------------------------------------------------------------------------
#if SODECL_SCHEDULE == 4 && (SODECL_SOLVER == 1 || SODECL_SOLVER == 2)
extern "C" __global__ void __launch_bounds__(SODECL_BLOCK) SODECL_KERNEL_ATTR
sodecl_team(double* __restrict__ t_state, double* __restrict__ y_state, const double* __restrict__ params,
            double* __restrict__ counters, double* __restrict__ traj, long long* __restrict__ newton,
            int orbits, int s0, int nseg, int final_snap, double* __restrict__ ckpt_t,
            double* __restrict__ ckpt_y, int ckpt_every, long long gstep0, double* __restrict__ stage_ws,
            long long ws_steps)
{
    SODECL_EXP_TABLE_INIT();
    const int gid = blockIdx.x * SODECL_BLOCK + threadIdx.x;
    const int i = gid / SODECL_GROUP;            // orbit: a team exits together
    const int lane = gid % SODECL_GROUP;
    if (i >= orbits) return;
    (void)counters;
    (void)newton;
    (void)stage_ws;
    (void)ws_steps;
    double y[_numeq_], yn[_numeq_], p[SODECL_NPAR_ALLOC];
    SODECL_UNROLL
    for (int m = 0; m < _numeq_; ++m) y[m] = y_state[(size_t)i * _numeq_ + m];
    for (int m = 0; m < _numpar_; ++m) p[m] = params[(size_t)i * _numpar_ + m];
    double t = t_state[i];
    const size_t slab = (size_t)_numeq_ * (size_t)orbits;
    long long ck = ckpt_y ? (gstep0 + ckpt_every - 1) / ckpt_every : 0;
    long long ck_step = ck * ckpt_every;
    long long g = gstep0;
    for (int s = 0; s < nseg; ++s) {
        if (traj) {
            double* o = traj + (size_t)(s0 + s) * slab + i;
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m)
                if (m % SODECL_GROUP == lane) o[(size_t)m * orbits] = y[m];
        }
        for (int it = 0; it < _numsteps_; ++it, ++g) {
            if (ckpt_y && g == ck_step) {
                const long long k = ck++;
                ck_step += ckpt_every;
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m)
                    if (m % SODECL_GROUP == lane) ckpt_y[(size_t)k * slab + (size_t)m * orbits + i] = y[m];
                if (lane == 0) ckpt_t[(size_t)k * orbits + i] = t;
            }
            sodecl_lane_step(t, y, yn, p);
            t = t + _m_dt_;
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) y[m] = yn[m];
        }
    }
    SODECL_UNROLL
    for (int m = 0; m < _numeq_; ++m) {
        if (m % SODECL_GROUP == lane) {
            if (traj && final_snap) traj[(size_t)(s0 + nseg) * slab + (size_t)m * orbits + i] = y[m];
            y_state[(size_t)i * _numeq_ + m] = y[m];
        }
    }
    if (lane == 0) t_state[i] = t;
}
#endif  // SODECL_SCHEDULE == 4

// ---------------------------------------------------------------------------
// Group schedule: SODECL_GROUP lanes per orbit (explicit solvers).
// Lane l of a group owns components l, l + G, l + 2G, ...; each stage vector
// is written to an LDS row, one wavefront barrier publishes it, and every lane then
// evaluates its components with the plugin's per-component entry point,
// reading the whole stage vector from LDS (same-address reads broadcast).
// Rows are double-buffered so one barrier per stage suffices.  (The group
// never spans wavefronts, so no barrier waits on another wavefront.)
// ---------------------------------------------------------------------------
#if SODECL_SCHEDULE == 2 && SODECL_HAS_COMPONENT && !SODECL_IS_IMPLICIT && (!SODECL_IS_SDE || SODECL_HAS_STOCH_COMPONENT)

#define SODECL_G SODECL_GROUP
#define SODECL_CPL ((_numeq_ + SODECL_G - 1) / SODECL_G)

// G is a power of two no larger than the wavefront, so the lanes of a group
// always sit in one wavefront and their exchanges through LDS need only
// wavefront-scope ordering.  Waves of a block run independently.
SODECL_INLINE void sodecl_group_sync() { sodecl_wave_sync(); }
#define SODECL_OPB (SODECL_BLOCK / SODECL_G)
#define SODECL_XSTRIDE (_numeq_ | 1)
#define SODECL_NSTAGE (SODECL_SOLVER == 2 ? 4 : 1)
// Stage workspace row of one step (adjoint): NSTAGE stage vectors of neq
// doubles, then the step's t.
#define SODECL_WROW (SODECL_NSTAGE * _numeq_ + 1)

// Drift of this lane's components at the published stage vector x.
//
// With the plugin's pairwise form (sode_pair, sode_system_pairsum: f_i is
// finished from s_i = sum_j pair(i, j) taken in j order, pair(i, j) ==
// -pair(j, i) exactly), every unordered pair is evaluated once: the pairs
// are (i, i + d mod N) for d = 1 .. (N-1)/2, and d = N/2 for the first half
// of the rows when N is even; the diagonal is +0.  In the two LDS layouts
// the pairs are spread over all G lanes of the group (G may exceed N, so a
// small system can use a whole wavefront per orbit); after one barrier each
// lane sums its own row in j order -- the same terms in the same order as
// the per-row loop, with half the evaluations.  Layouts (SODECL_USE_PAIRS):
//   1  the full matrix S[N][N+1]: the evaluating lane writes the value to
//      S[i][j] and its negation to S[j][i]; the row sum reads S[i][j];
//   2  only the evaluated values, V[d-1][i] = pair(i, i + d mod N), half the
//      storage: the row sum reads V[(j-i) mod N - 1][i] when that offset is
//      row i's own, else -V[(i-j) mod N - 1][j].
//   3  no storage (one component per lane, G >= N): at offset d the lane of
//      row i evaluates v = pair(i, i + d mod N), adds v and subtracts the v
//      of the lane of row i - d mod N, whose pair (i - d, i) is the negation
//      of row i's term j = i - d; the lane exchange is one bpermute per
//      32-bit half.  Each row sums the same N - 1 terms, in the order of the
//      offsets rather than of j, so the drift agrees with the per-row form
//      to rounding of the sum (not bit for bit).  Without LDS matrices the
//      kernel keeps the per-component form's occupancy.
// Otherwise each component comes from sode_system_i.
#if SODECL_USE_PAIRS
#define SODECL_HALF ((_numeq_ - 1) / 2)
#define SODECL_EVEN ((_numeq_ % 2) == 0)
#if SODECL_USE_PAIRS == 3
static_assert(SODECL_CPL == 1, "the lane-exchange pairwise drift needs one component per lane");
#define SODECL_PAIR_LDS 1
SODECL_INLINE void sodecl_group_eval(double t, double* x, double* pp, double* S, int lane, double (&f)[SODECL_CPL])
{
    (void)S;
    const int i = lane;
    const bool row = i < _numeq_;
    const int ir = row ? i : 0;   // lanes past the last row index row 0 and discard the result
    double acc = 0.0;
    for (int d = 1; d <= SODECL_HALF; ++d) {
        int j, src;
        if ((_numeq_ & (_numeq_ - 1)) == 0) {   // N a power of two: wrap by mask
            j = (ir + d) & (_numeq_ - 1);
            src = (ir - d) & (_numeq_ - 1);
        } else {
            j = ir + d;
            if (j >= _numeq_) j -= _numeq_;
            src = ir - d;
            if (src < 0) src += _numeq_;
        }
        const double v = row ? sode_pair(i, j, t, x, pp) : 0.0;
        double w;
        if (_numeq_ == SODECL_G && (SODECL_G & (SODECL_G - 1)) == 0) {
            // ds_bpermute selects the source lane by address bits 7..2, so
            // (i - d) * 4 wraps around a whole wavefront by itself; a smaller
            // group keeps the wrapped offset inside its own lanes
            int addr = (i - d) * 4;
            if (SODECL_G < 64) addr = (addr & (4 * SODECL_G - 1)) | (int)((threadIdx.x & (64 - SODECL_G)) * 4);
            const long long bits = __double_as_longlong(v);
            const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(bits & 0xffffffffll));
            const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(bits >> 32));
            w = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
        } else {
            w = __shfl(v, src, SODECL_G);
        }
        acc = acc + v;
        acc = acc - w;
    }
    if (SODECL_EVEN) {
        // offset N/2: rows below N/2 evaluate, rows from N/2 up take the
        // negation from row i - N/2
        const int h = _numeq_ / 2;
        const double v = (i < h) ? sode_pair(i, i + h, t, x, pp) : 0.0;
        const double w = __shfl(v, (row && i >= h) ? i - h : 0, SODECL_G);
        acc = i < h ? acc + v : acc - w;
    }
    f[0] = row ? sode_system_pairsum(i, t, x, pp, acc) : 0.0;
}
#else
#if SODECL_USE_PAIRS == 1
#define SODECL_SSTRIDE (_numeq_ + 1)
#define SODECL_PAIR_LDS (_numeq_ * SODECL_SSTRIDE)
#else
#define SODECL_PAIR_LDS (_numeq_ * (SODECL_HALF + (SODECL_EVEN ? 1 : 0)))
#endif
// Unordered pairs in a flattened order, k = (d - 1) N + i for the pair
// (i, i + d mod N): offsets d = 1 .. HALF for every row, then d = N/2 for
// rows i < N/2 when N is even.
#define SODECL_NPAIRS (_numeq_ * SODECL_HALF + (SODECL_EVEN ? _numeq_ / 2 : 0))
SODECL_INLINE void sodecl_group_eval(double t, double* x, double* pp, double* S, int lane, double (&f)[SODECL_CPL])
{
    // (the diagonal of the full matrix is written once per launch, by
    // sodecl_pairs_init)
    // the pairs go round-robin to all G lanes of the group (lanes past the
    // last row evaluate pairs too), at most ceil(NPAIRS / G) each
    for (int k = lane; k < SODECL_NPAIRS; k += SODECL_G) {
        const int d = k / _numeq_ + 1;
        const int i = k - (d - 1) * _numeq_;
        int j = i + d;
        if (j >= _numeq_) j -= _numeq_;
        const double v = sode_pair(i, j, t, x, pp);
#if SODECL_USE_PAIRS == 1
        S[i * SODECL_SSTRIDE + j] = v;
        S[j * SODECL_SSTRIDE + i] = -v;
#else
        S[k] = v;
#endif
    }
    sodecl_group_sync();
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int i = lane + c * SODECL_G;
        f[c] = 0.0;
        if (i < _numeq_) {
            double acc = 0;
#if SODECL_USE_PAIRS == 1
            for (int j = 0; j < _numeq_; ++j) acc = acc + S[i * SODECL_SSTRIDE + j];
#else
            // offsets row i owns: 1 .. HALF, and N/2 for i < N/2 when N is even
            const int own_max = SODECL_HALF + ((SODECL_EVEN && i < _numeq_ / 2) ? 1 : 0);
            int dd = i == 0 ? 0 : _numeq_ - i;   // (j - i) mod N at j = 0
            for (int j = 0; j < _numeq_; ++j) {
                const bool own = dd <= own_max;
                // dd == 0 (the diagonal) reads a valid slot and discards it
                const int k = dd == 0 ? i : own ? (dd - 1) * _numeq_ + i : (_numeq_ - dd - 1) * _numeq_ + j;
                const double v = dd == 0 ? 0.0 : S[k];
                acc = acc + (own ? v : -v);
                dd = dd + 1 == _numeq_ ? 0 : dd + 1;
            }
#endif
            f[c] = sode_system_pairsum(i, t, x, pp, acc);
        }
    }
}
#endif
#define SODECL_EVAL(T, X, F) sodecl_group_eval(T, X, pp, S, lane, F)
#else
SODECL_INLINE void sodecl_group_eval(double t, double* x, double* pp, int lane, double (&f)[SODECL_CPL])
{
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int comp = lane + c * SODECL_G;
        f[c] = comp < _numeq_ ? sode_system_i(comp, t, x, pp) : 0.0;
    }
}
#define SODECL_EVAL(T, X, F) sodecl_group_eval(T, X, pp, lane, F)
#endif

// Once per launch, before the first evaluation: the full pair matrix's
// diagonal.  pair(i, i) = -pair(i, i) is zero for finite states (and a
// non-finite state makes the whole row non-finite), so it is never evaluated,
// and no evaluation writes the diagonal.
SODECL_INLINE void sodecl_pairs_init(double* S, int lane)
{
#if SODECL_USE_PAIRS == 1
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int i = lane + c * SODECL_G;
        if (i < _numeq_) S[i * SODECL_SSTRIDE + i] = 0.0;
    }
#else
    (void)S;
    (void)lane;
#endif
}

// Writes this lane's components of v into LDS row x.
SODECL_INLINE void sodecl_publish(double* x, const double (&v)[SODECL_CPL], int lane)
{
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int comp = lane + c * SODECL_G;
        if (comp < _numeq_) x[comp] = v[c];
    }
}

// Optional copy of a stage vector to a per-orbit workspace row (adjoint).
SODECL_INLINE void sodecl_store_row(double* row, const double (&v)[SODECL_CPL], int lane)
{
    if (!row) return;
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int comp = lane + c * SODECL_G;
        if (comp < _numeq_) row[comp] = v[c];
    }
}

#if SODECL_SOLVER == 1
// Euler (euler.cl:10-18).  stages: null, or one row receiving y.
SODECL_INLINE void sodecl_group_step(double t, double (&y)[SODECL_CPL], double (*xs)[SODECL_OPB][SODECL_XSTRIDE],
                                     int& buf, int slot, int lane, double* pp, double* S, double* stages)
{
    double f[SODECL_CPL];
    sodecl_store_row(stages, y, lane);
    sodecl_publish(xs[buf][slot], y, lane);
    sodecl_group_sync();
    SODECL_EVAL(t, xs[buf][slot], f);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c)
        if (lane + c * SODECL_G < _numeq_) y[c] = y[c] + f[c] * _m_dt_;
    buf ^= 1;
}
#elif SODECL_SOLVER == 2
// Classic RK4 (rk4.cl:17-64), one LDS stage row per evaluation.  stages:
// null, or four rows receiving x1 = y, x2, x3, x4.
SODECL_INLINE void sodecl_group_step(double t, double (&y)[SODECL_CPL], double (*xs)[SODECL_OPB][SODECL_XSTRIDE],
                                     int& buf, int slot, int lane, double* pp, double* S, double* stages)
{
    double acc[SODECL_CPL], xn[SODECL_CPL], f[SODECL_CPL];
    sodecl_store_row(stages, y, lane);
    sodecl_publish(xs[buf][slot], y, lane);
    sodecl_group_sync();
    SODECL_EVAL(t, xs[buf][slot], f);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        acc[c] = f[c];
        xn[c] = y[c] + (_m_dt_ * 0.50) * f[c];
    }
    buf ^= 1;
    sodecl_store_row(stages ? stages + _numeq_ : nullptr, xn, lane);
    sodecl_publish(xs[buf][slot], xn, lane);
    sodecl_group_sync();
    SODECL_EVAL(t + 0.50 * _m_dt_, xs[buf][slot], f);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        acc[c] = acc[c] + 2.00 * f[c];
        xn[c] = y[c] + (_m_dt_ * 0.50) * f[c];
    }
    buf ^= 1;
    sodecl_store_row(stages ? stages + 2 * _numeq_ : nullptr, xn, lane);
    sodecl_publish(xs[buf][slot], xn, lane);
    sodecl_group_sync();
    SODECL_EVAL(t + 0.50 * _m_dt_, xs[buf][slot], f);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        acc[c] = acc[c] + 2.00 * f[c];
        xn[c] = y[c] + _m_dt_ * f[c];
    }
    buf ^= 1;
    sodecl_store_row(stages ? stages + 3 * _numeq_ : nullptr, xn, lane);
    sodecl_publish(xs[buf][slot], xn, lane);
    sodecl_group_sync();
    SODECL_EVAL(t + _m_dt_, xs[buf][slot], f);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c)
        if (lane + c * SODECL_G < _numeq_) y[c] = y[c] + SODECL_RK4_DIV6((acc[c] + f[c]) * _m_dt_);
    buf ^= 1;
}
#endif

// Group-schedule kernel.  Arguments as sodecl_lane, plus optional
// checkpoints: when ckpt_y is not null, the state (and t) before every global
// step g with g % ckpt_every == 0 is written to ckpt_y[g/ckpt_every][neq][orbits]
// and ckpt_t[g/ckpt_every][orbits]; gstep0 is the global index of this
// launch's first step.
extern "C" __global__ void __launch_bounds__(SODECL_BLOCK) SODECL_KERNEL_ATTR
sodecl_group(double* __restrict__ t_state, double* __restrict__ y_state, const double* __restrict__ params,
             double* __restrict__ counters, double* __restrict__ traj, long long* __restrict__ newton,
             int orbits, int s0, int nseg, int final_snap, double* __restrict__ ckpt_t,
             double* __restrict__ ckpt_y, int ckpt_every, long long gstep0, double* __restrict__ stage_ws,
             long long ws_steps)
{
    SODECL_EXP_TABLE_INIT();
    __shared__ double sh_x[2][SODECL_OPB][SODECL_XSTRIDE];
    __shared__ double sh_p[SODECL_OPB][SODECL_NPAR_ALLOC];
#if SODECL_IS_SDE
    __shared__ double sh_n[2][SODECL_OPB][SODECL_NNOI_ALLOC];
#endif
#if SODECL_USE_PAIRS == 1 || SODECL_USE_PAIRS == 2
    __shared__ double sh_S[SODECL_OPB][SODECL_PAIR_LDS];
#endif
    const int lane = threadIdx.x % SODECL_G;
    const int slot = threadIdx.x / SODECL_G;
    const int orbit = blockIdx.x * SODECL_OPB + slot;
    const bool active = orbit < orbits;
    (void)newton;
#if SODECL_USE_PAIRS == 1 || SODECL_USE_PAIRS == 2
    double* const S = sh_S[slot];
#else
    double* const S = nullptr;
#endif
    (void)S;
    sodecl_pairs_init(S, lane);
#if SODECL_IS_SDE
    (void)stage_ws;   // no adjoint of Euler-Maruyama: never stored
    (void)ws_steps;
#endif

    for (int k = lane; k < _numpar_; k += SODECL_G)
        sh_p[slot][k] = active ? params[(size_t)orbit * _numpar_ + k] : 0.0;
    double y[SODECL_CPL];
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int comp = lane + c * SODECL_G;
        y[c] = (active && comp < _numeq_) ? y_state[(size_t)orbit * _numeq_ + comp] : 0.0;
    }
    double t = active ? t_state[orbit] : 0.0;
    double* const pp = sh_p[slot];
    const size_t slab = (size_t)_numeq_ * (size_t)orbits;
#if SODECL_IS_SDE
    double cnt = active ? counters[orbit] : 0.0;
    const double dtsqrt = sqrt(_m_dt_);
    const int calls = (int)ceilf((float)((float)(_numnoi_) * 0.5));
    bool noise_ready = false;   // this step's noise was drawn with the previous step's
#endif
    int buf = 0;
    long long g = gstep0;
    // checkpoint k is the state before global step k * ckpt_every
    long long ck = ckpt_y ? (gstep0 + ckpt_every - 1) / ckpt_every : 0;
    long long ck_step = ck * ckpt_every;

    for (int s = 0; s < nseg; ++s) {
        if (traj && active) {
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                const int comp = lane + c * SODECL_G;
                if (comp < _numeq_) traj[(size_t)(s0 + s) * slab + (size_t)comp * orbits + orbit] = y[c];
            }
        }
        // The segment's steps in runs that end at the next checkpoint, so that
        // the step loop tests only its count; the stage-vector store is chosen
        // once per run (a kernel argument, the same for every lane, so the
        // step's lane exchanges see the whole wavefront either way).  With one
        // wavefront per SIMD each scalar instruction in the step loop takes an
        // issue cycle from the VALU.
        for (int it = 0; it < _numsteps_;) {
            int run = _numsteps_ - it;
            if (ckpt_y) {
                if (g == ck_step) {
                    const long long k = ck++;
                    ck_step += ckpt_every;
                    if (active) {
#pragma unroll
                        for (int c = 0; c < SODECL_CPL; ++c) {
                            const int comp = lane + c * SODECL_G;
                            if (comp < _numeq_) ckpt_y[(size_t)k * slab + (size_t)comp * orbits + orbit] = y[c];
                        }
                        if (lane == 0) ckpt_t[(size_t)k * orbits + orbit] = t;
                    }
                }
                if (ck_step - g < run) run = (int)(ck_step - g);   // g < ck_step here
            }
            const int end = it + run;
#if SODECL_IS_SDE
            for (; it < end; ++it, ++g) {
                // Noise: Threefry call k of this step goes to lane k of the group;
                // the counter of call k is counter + k, as the reference's
                // sequential +1 per call (stochastic_solver_caller.cl:44-66).
                // The noise does not depend on the state: a group with at least
                // twice as many lanes as calls draws the next step of the same
                // segment in the same pass (lanes calls .. 2 calls - 1, counters
                // counter + calls + k, into the next step's buffer), and that
                // step draws none.
                if (!noise_ready) {
                    const int n = (SODECL_G >= 2 * SODECL_CALLS && it + 1 < _numsteps_) ? 2 * calls : calls;
                    for (int k = lane; k < n; k += SODECL_G) {
                        const int nb = k < calls ? buf : (buf ^ 1);
                        const int kk = k < calls ? k : k - calls;
                        double z0, z1;
                        sodecl_noise_pair(orbit, cnt + (double)k, dtsqrt, &z0, &z1);
                        if (2 * kk < _numnoi_) sh_n[nb][slot][2 * kk] = z0;
                        if (2 * kk + 1 < _numnoi_) sh_n[nb][slot][2 * kk + 1] = z1;
                    }
                    noise_ready = n > calls;
                } else {
                    noise_ready = false;
                }
                cnt = cnt + (double)calls;
                sodecl_publish(sh_x[buf][slot], y, lane);
                sodecl_group_sync();
                // Euler-Maruyama (stochastic_euler.cl:10-22)
                {
                    double f[SODECL_CPL];
                    SODECL_EVAL(t, sh_x[buf][slot], f);
#pragma unroll
                    for (int c = 0; c < SODECL_CPL; ++c) {
                        const int comp = lane + c * SODECL_G;
                        if (comp < _numeq_) {
                            const double gn = sode_system_stoch_i(comp, t, sh_x[buf][slot], pp, sh_n[buf][slot]);
                            y[c] = y[c] + f[c] * _m_dt_ + gn;
                        }
                    }
                }
                buf ^= 1;
                t = t + _m_dt_;
            }
#else
            if (stage_ws) {
                for (; it < end; ++it, ++g) {
                    double* row = nullptr;
                    if (active) {
                        // every step's stage vectors for an adjoint without recomputation
                        row = stage_ws + ((size_t)orbit * (size_t)ws_steps + (size_t)g) * SODECL_WROW;
                        if (lane == 0) row[SODECL_NSTAGE * _numeq_] = t;
                    }
                    sodecl_group_step(t, y, sh_x, buf, slot, lane, pp, S, row);
                    t = t + _m_dt_;
                }
            } else {
                for (; it < end; ++it, ++g) {
                    sodecl_group_step(t, y, sh_x, buf, slot, lane, pp, S, nullptr);
                    t = t + _m_dt_;
                }
            }
#endif
        }
    }
    if (active) {
#pragma unroll
        for (int c = 0; c < SODECL_CPL; ++c) {
            const int comp = lane + c * SODECL_G;
            if (comp < _numeq_) {
                if (traj && final_snap)
                    traj[(size_t)(s0 + nseg) * slab + (size_t)comp * orbits + orbit] = y[c];
                y_state[(size_t)orbit * _numeq_ + comp] = y[c];
            }
        }
        if (lane == 0) {
            t_state[orbit] = t;
#if SODECL_IS_SDE
            counters[orbit] = cnt;
#endif
        }
    }
}
#endif  // group schedule
