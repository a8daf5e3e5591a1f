// Discrete adjoint of the explicit solvers under the group schedule.
//
// Gradients of L(trajectory) with respect to each orbit's initial state and
// parameters, given dL/d(snapshot) for every snapshot.  The adjoint is the
// exact transpose of the forward step as the forward kernel computes it
// (classic RK4 of rk4.cl:17-64 or Euler of euler.cl:10-18), so it matches a
// forward-sensitivity computation to rounding.
//
// Memory is bounded by checkpointing (Griewank's revolve with one level):
// the forward kernel stores the state every ckpt_every steps (sqrt of the
// step count by default); here each segment is recomputed from its
// checkpoint, with the stage vectors of every step written to a per-orbit
// workspace, and then swept backwards.  Where memory allows (an MI355X holds
// 288 GB), the forward kernel can instead write every step's stage vectors
// to a whole-run workspace (stored = 1, one segment): the sweep then needs
// no recomputation.  The plugin supplies the
// vector-Jacobian product of its drift, evaluated cooperatively by the group:
//
//   void sode_system_vjp_group(double t, double x[], double lam[], double p[],
//                              double gx[], double gp[])
//     gx[i] = sum_k lam[k] df_k/dx_i for this lane's components
//             i = sodecl_lane() + c * sodecl_group_size();
//     gp[k] += sum_i lam[i] df_i/dp_k, every entry updated by one lane only.
#if SODECL_SCHEDULE == 2 && defined(SODECL_ADJOINT) && SODECL_HAS_VJP_GROUP && \
    (SODECL_SOLVER == 1 || SODECL_SOLVER == 2)

// One VJP through LDS: publish point v and cotangent l, barrier, evaluate.
SODECL_INLINE void sodecl_group_vjp(double t, const double (&v)[SODECL_CPL], const double (&l)[SODECL_CPL],
                                    double (&g)[SODECL_CPL], double (*sv)[SODECL_OPB][SODECL_XSTRIDE],
                                    double (*sl)[SODECL_OPB][SODECL_XSTRIDE], double* sg, int& vb, int slot,
                                    int lane, double* pp, double* gp)
{
    sodecl_publish(sv[vb][slot], v, lane);
    sodecl_publish(sl[vb][slot], l, lane);
    sodecl_group_sync();
    sode_system_vjp_group(t, sv[vb][slot], sl[vb][slot], pp, sg, gp);
#pragma unroll
    for (int c = 0; c < SODECL_CPL; ++c) {
        const int comp = lane + c * SODECL_G;
        g[c] = comp < _numeq_ ? sg[comp] : 0.0;
    }
    vb ^= 1;
}

// Per-orbit workspace: one row of SODECL_WROW doubles (NSTAGE stage vectors
// of neq doubles, then the step's t) per step of a segment.

extern "C" __global__ void __launch_bounds__(SODECL_BLOCK)
sodecl_group_adjoint(const double* __restrict__ ckpt_t, const double* __restrict__ ckpt_y,
                     const double* __restrict__ params, const double* __restrict__ cot,
                     double* __restrict__ grad_y0, double* __restrict__ grad_p, double* __restrict__ work,
                     int orbits, int total_steps, int ckpt_every, int stored)
{
    SODECL_EXP_TABLE_INIT();
    __shared__ double sh_x[2][SODECL_OPB][SODECL_XSTRIDE];
    __shared__ double sh_v[2][SODECL_OPB][SODECL_XSTRIDE];
    __shared__ double sh_l[2][SODECL_OPB][SODECL_XSTRIDE];
    __shared__ double sh_g[SODECL_OPB][SODECL_XSTRIDE];
    __shared__ double sh_p[SODECL_OPB][SODECL_NPAR_ALLOC];
    __shared__ double sh_gp[SODECL_OPB][SODECL_NPAR_ALLOC];
    const int lane = threadIdx.x % SODECL_G;
    const int slot = threadIdx.x / SODECL_G;
    const int orbit = blockIdx.x * SODECL_OPB + slot;
    const bool active = orbit < orbits;
    const int ob = active ? orbit : 0;   // inactive groups read orbit 0's inputs and store no result
    double* const pp = sh_p[slot];
    double* const gp = sh_gp[slot];
#if SODECL_USE_PAIRS == 1 || SODECL_USE_PAIRS == 2
    __shared__ double sh_S[SODECL_OPB][SODECL_PAIR_LDS];
    double* const S = sh_S[slot];
#else
    double* const S = nullptr;
#endif
    sodecl_pairs_init(S, lane);
    for (int k = lane; k < _numpar_; k += SODECL_G) {
        pp[k] = params[(size_t)ob * _numpar_ + k];
        gp[k] = 0.0;
    }
    const size_t slab = (size_t)_numeq_ * (size_t)orbits;
    double lam[SODECL_CPL];
    // cotangent of the final snapshot (step total_steps)
    {
        const long long last = total_steps / _numsteps_;
#pragma unroll
        for (int c = 0; c < SODECL_CPL; ++c) {
            const int comp = lane + c * SODECL_G;
            lam[c] = comp < _numeq_ ? cot[(size_t)last * slab + (size_t)comp * orbits + ob] : 0.0;
        }
    }
    // the workspace is sized for every group of the grid, active or not
    double* const wk = work + (size_t)orbit * ckpt_every * SODECL_WROW;
    const int nseg = (total_steps + ckpt_every - 1) / ckpt_every;
    int buf = 0, vb = 0;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int sb = seg * ckpt_every;
        const int se = min(total_steps, sb + ckpt_every);
        // forward recompute of the segment, same arithmetic as the forward
        // kernel (skipped when the forward kernel stored every step's rows)
        if (!stored) {
            double y[SODECL_CPL];
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                const int comp = lane + c * SODECL_G;
                y[c] = comp < _numeq_ ? ckpt_y[(size_t)seg * slab + (size_t)comp * orbits + ob] : 0.0;
            }
            double t = ckpt_t[(size_t)seg * orbits + ob];
            for (int s = sb; s < se; ++s) {
                double* row = wk + (size_t)(s - sb) * SODECL_WROW;
                if (lane == 0) row[SODECL_NSTAGE * _numeq_] = t;
                sodecl_group_step(t, y, sh_x, buf, slot, lane, pp, S, row);
                t = t + _m_dt_;
            }
        }
        __syncthreads();   // workspace rows written by other lanes are read below
        // backward sweep
        for (int s = se - 1; s >= sb; --s) {
            const double* row = wk + (size_t)(s - sb) * SODECL_WROW;
            const double ts = row[SODECL_NSTAGE * _numeq_];
            double x[SODECL_CPL], gv[SODECL_CPL];
#if SODECL_SOLVER == 1
            // y' = y + f(t, y) dt: lam_s = lam' + J^T (dt lam'), gp += f_p^T (dt lam')
            double l1[SODECL_CPL];
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                const int comp = lane + c * SODECL_G;
                x[c] = comp < _numeq_ ? row[comp] : 0.0;
                l1[c] = _m_dt_ * lam[c];
            }
            sodecl_group_vjp(ts, x, l1, gv, sh_v, sh_l, sh_g[slot], vb, slot, lane, pp, gp);
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) lam[c] = lam[c] + gv[c];
#else
            // y' = y + (k1 + 2 k2 + 2 k3 + k4) dt / 6, k_i = f(t_i, x_i),
            // x2 = y + dt/2 k1, x3 = y + dt/2 k2, x4 = y + dt k3
            double kb1[SODECL_CPL], kb2[SODECL_CPL], kb3[SODECL_CPL], kb4[SODECL_CPL], yb[SODECL_CPL];
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                kb1[c] = (_m_dt_ / 6.00) * lam[c];
                kb2[c] = (2.00 * _m_dt_ / 6.00) * lam[c];
                kb3[c] = kb2[c];
                kb4[c] = kb1[c];
                yb[c] = lam[c];
            }
#define SODECL_LOAD_STAGE(S)                                                   \
    _Pragma("unroll") for (int c = 0; c < SODECL_CPL; ++c) {                   \
        const int comp = lane + c * SODECL_G;                                  \
        x[c] = comp < _numeq_ ? row[(S) * _numeq_ + comp] : 0.0;               \
    }
            SODECL_LOAD_STAGE(3)
            sodecl_group_vjp(ts + _m_dt_, x, kb4, gv, sh_v, sh_l, sh_g[slot], vb, slot, lane, pp, gp);
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                yb[c] += gv[c];
                kb3[c] += _m_dt_ * gv[c];
            }
            SODECL_LOAD_STAGE(2)
            sodecl_group_vjp(ts + 0.50 * _m_dt_, x, kb3, gv, sh_v, sh_l, sh_g[slot], vb, slot, lane, pp, gp);
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                yb[c] += gv[c];
                kb2[c] += (_m_dt_ * 0.50) * gv[c];
            }
            SODECL_LOAD_STAGE(1)
            sodecl_group_vjp(ts + 0.50 * _m_dt_, x, kb2, gv, sh_v, sh_l, sh_g[slot], vb, slot, lane, pp, gp);
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) {
                yb[c] += gv[c];
                kb1[c] += (_m_dt_ * 0.50) * gv[c];
            }
            SODECL_LOAD_STAGE(0)
            sodecl_group_vjp(ts, x, kb1, gv, sh_v, sh_l, sh_g[slot], vb, slot, lane, pp, gp);
#pragma unroll
            for (int c = 0; c < SODECL_CPL; ++c) lam[c] = yb[c] + gv[c];
#undef SODECL_LOAD_STAGE
#endif
            // snapshot of step s: add its cotangent
            if (s % _numsteps_ == 0) {
                const long long j = s / _numsteps_;
#pragma unroll
                for (int c = 0; c < SODECL_CPL; ++c) {
                    const int comp = lane + c * SODECL_G;
                    if (comp < _numeq_) lam[c] += cot[(size_t)j * slab + (size_t)comp * orbits + ob];
                }
            }
        }
        __syncthreads();   // the next segment overwrites the workspace
    }
    if (active) {
#pragma unroll
        for (int c = 0; c < SODECL_CPL; ++c) {
            const int comp = lane + c * SODECL_G;
            if (comp < _numeq_) grad_y0[(size_t)orbit * _numeq_ + comp] = lam[c];
        }
    }
    __syncthreads();
    if (active)
        for (int k = lane; k < _numpar_; k += SODECL_G) grad_p[(size_t)orbit * _numpar_ + k] = gp[k];
}
#endif  // adjoint

// ---------------------------------------------------------------------------
// Discrete adjoint of the implicit solvers, lane schedule (one orbit per
// lane).  A step maps y to the converged Newton solution g of
//   implicit Euler      F = g - y - dt f(t, g)               (ie.cl:50-117)
//   implicit midpoint   F = g - y - dt f(t, m), m = (y+g)/2  (im.cl:43-118)
// and its derivative is that of the implicit equation: with
// A = dF/dg = I - c dt J (c = 1 at g, c = 1/2 at m) the cotangent lam of g
// gives mu = A^-T lam and
//   dL/dy += mu                     (IE)   mu + (dt/2) J^T mu   (IM)
//   dL/dp += dt (df/dp)^T mu        at g (IE) or m (IM).
// A^T mu = lam is solved like the forward Newton system, by sweeps of a
// tridiagonal (Thomas) solve: mu += T^-1 (lam - A^T mu) with T the
// tridiagonal band of A^T, until the residual's L1 norm is below 1e-15 of
// lam's or after SODECL_NEWTON_MAXIT sweeps.  (The forward's sweep keeps the
// reference's index quirks because its iterates must match the reference;
// the adjoint has no reference counterpart and uses the band of A^T as it
// is.)  For a tridiagonal J one sweep is exact.
//
// The plugin provides df/dy as calc_jacobian_exact(y, jac[neq][neq], p), or
// when that is absent as calc_jacobian (the Newton matrix's source, which a
// plugin may define differently: the reference's oculomotor example has one
// entry without its sign), and calc_jacobian_params(y, jp[neq][npar], p)
// (df/dp).  Checkpoint
// segments are recomputed with the forward Newton step; the states of a
// segment go to the workspace laid out [row][neq + 1][orbits] (state and t
// of row k = step sb + k, rows 0..ckpt_every), orbit-contiguous so the
// lanes' accesses coalesce.
// ---------------------------------------------------------------------------
#if SODECL_SCHEDULE == 1 && defined(SODECL_ADJOINT) && SODECL_IS_IMPLICIT && SODECL_HAS_JACOBIAN && \
    SODECL_HAS_JACOBIAN_PARAMS

#define SODECL_IMPLICIT_C (SODECL_SOLVER == 3 ? 1.0 : 0.50)

// Tridiagonal solve (Thomas): row m is sub[m] x[m-1] + dia[m] x[m] +
// sup[m] x[m+1] = rhs[m].
SODECL_INLINE void sodecl_tridiag(const double* sub, const double* dia, const double* sup, const double* rhs,
                                  double* x)
{
    double cp[_numeq_], dp[_numeq_];
    cp[0] = sup[0] / dia[0];
    dp[0] = rhs[0] / dia[0];
    SODECL_UNROLL
    for (int m = 1; m < _numeq_; ++m) {
        const double den = dia[m] - sub[m] * cp[m - 1];
        cp[m] = sup[m] / den;
        dp[m] = (rhs[m] - sub[m] * dp[m - 1]) / den;
    }
    x[_numeq_ - 1] = dp[_numeq_ - 1];
    SODECL_UNROLL
    for (int m = _numeq_ - 2; m >= 0; --m) x[m] = dp[m] - cp[m] * x[m + 1];
}

extern "C" __global__ void __launch_bounds__(SODECL_BLOCK) SODECL_KERNEL_ATTR
sodecl_lane_adjoint(const double* __restrict__ ckpt_t, const double* __restrict__ ckpt_y,
                    const double* __restrict__ params, const double* __restrict__ cot,
                    double* __restrict__ grad_y0, double* __restrict__ grad_p, double* __restrict__ work,
                    int orbits, int total_steps, int ckpt_every, int stored)
{
    SODECL_EXP_TABLE_INIT();
    (void)stored;   // always recomputes: stored rows exist for the explicit solvers only
    const int i = blockIdx.x * SODECL_BLOCK + threadIdx.x;
    if (i >= orbits) return;
    double p[SODECL_NPAR_ALLOC], gp[SODECL_NPAR_ALLOC];
    for (int m = 0; m < _numpar_; ++m) {
        p[m] = params[(size_t)i * _numpar_ + m];
        gp[m] = 0.0;
    }
    const size_t slab = (size_t)_numeq_ * (size_t)orbits;
    const size_t wrow = (size_t)(_numeq_ + 1) * (size_t)orbits;   // one workspace row
    double lam[_numeq_];
    {
        const long long last = total_steps / _numsteps_;
        SODECL_UNROLL
        for (int m = 0; m < _numeq_; ++m) lam[m] = cot[(size_t)last * slab + (size_t)m * orbits + i];
    }
    const int nseg = (total_steps + ckpt_every - 1) / ckpt_every;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int sb = seg * ckpt_every;
        const int se = min(total_steps, sb + ckpt_every);
        double y[_numeq_], g[_numeq_];
        SODECL_UNROLL
        for (int m = 0; m < _numeq_; ++m) y[m] = ckpt_y[(size_t)seg * slab + (size_t)m * orbits + i];
        double t = ckpt_t[(size_t)seg * orbits + i];
        // forward recompute of the segment with the forward Newton step
        for (int s = sb; s <= se; ++s) {
            double* row = work + (size_t)(s - sb) * wrow + i;
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) row[(size_t)m * orbits] = y[m];
            row[(size_t)_numeq_ * orbits] = t;
            if (s == se) break;
            sodecl_implicit_step(t, y, g, p);
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) y[m] = g[m];
            t = t + _m_dt_;
        }
        // backward sweep
        for (int s = se - 1; s >= sb; --s) {
            const double* r0 = work + (size_t)(s - sb) * wrow + i;
            const double* r1 = r0 + wrow;
            double x[_numeq_];   // point of J and df/dp: g (IE) or (y + g) / 2 (IM)
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) {
#if SODECL_SOLVER == 3
                x[m] = r1[(size_t)m * orbits];
#else
                x[m] = (r0[(size_t)m * orbits] + r1[(size_t)m * orbits]) * 0.5;
#endif
            }
            double jac[_numeq_][_numeq_];
#if SODECL_HAS_JACOBIAN_EXACT
            calc_jacobian_exact(x, jac, p);
#else
            calc_jacobian(x, jac, p);
#endif
            // tridiagonal band of A^T: A^T[m][m-1] = A[m-1][m], A^T[m][m+1] = A[m+1][m]
            const double cdt = SODECL_IMPLICIT_C * _m_dt_;
            double tsub[_numeq_], di[_numeq_], tsup[_numeq_];
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) {
                di[m] = 1.0 - cdt * jac[m][m];
                tsub[m] = m > 0 ? -cdt * jac[m - 1][m] : 0.0;
                tsup[m] = m + 1 < _numeq_ ? -cdt * jac[m + 1][m] : 0.0;
            }
            // A^T mu = lam, A = I - c dt J
            double mu[_numeq_], res[_numeq_], d[_numeq_];
            double lnorm = 0.0;
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) {
                mu[m] = 0.0;
                res[m] = lam[m];
                lnorm += fabs(lam[m]);
            }
            for (int it = 0; it < SODECL_NEWTON_MAXIT; ++it) {
                double rnorm = 0.0;
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m) rnorm += fabs(res[m]);
                if (!(rnorm > 1e-15 * lnorm)) break;
                sodecl_tridiag(tsub, di, tsup, res, d);
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m) mu[m] += d[m];
                // res = lam - A^T mu = lam - mu + c dt J^T mu
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m) {
                    double a = 0.0;
                    SODECL_UNROLL
                    for (int k = 0; k < _numeq_; ++k) a += jac[k][m] * mu[k];
                    res[m] = lam[m] - mu[m] + cdt * a;
                }
            }
            // dL/dp += dt (df/dp)^T mu
            if (_numpar_ > 0) {
                double jp[_numeq_][SODECL_NPAR_ALLOC];
                calc_jacobian_params(x, jp, p);
                for (int k = 0; k < _numpar_; ++k) {
                    double a = 0.0;
                    SODECL_UNROLL
                    for (int m = 0; m < _numeq_; ++m) a += jp[m][k] * mu[m];
                    gp[k] += _m_dt_ * a;
                }
            }
            // dL/dy
            SODECL_UNROLL
            for (int m = 0; m < _numeq_; ++m) {
#if SODECL_SOLVER == 3
                lam[m] = mu[m];
#else
                double a = 0.0;
                SODECL_UNROLL
                for (int k = 0; k < _numeq_; ++k) a += jac[k][m] * mu[k];
                lam[m] = mu[m] + (0.50 * _m_dt_) * a;
#endif
            }
            if (s % _numsteps_ == 0) {
                const long long j = s / _numsteps_;
                SODECL_UNROLL
                for (int m = 0; m < _numeq_; ++m) lam[m] += cot[(size_t)j * slab + (size_t)m * orbits + i];
            }
        }
    }
    SODECL_UNROLL
    for (int m = 0; m < _numeq_; ++m) grad_y0[(size_t)i * _numeq_ + m] = lam[m];
    for (int k = 0; k < _numpar_; ++k) grad_p[(size_t)i * _numpar_ + k] = gp[k];
}
#endif  // implicit adjoint
