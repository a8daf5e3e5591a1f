// Per-size-class part of the batched implicit integrator.  The host emits
// this text once per system size class, inside a namespace, after defining
// _numeq_ for that class and including the system plugin there.
//
// One orbit is integrated by one wavefront (SODECL_BG = 64 lanes):
//   * components are owned strided (comp = lane + 64 c) for the right-hand
//     side, the band Jacobian and the Newton update;
//   * the tridiagonal solve of each Newton iteration runs in parallel over
//     contiguous row partitions (lane k owns rows [kR, kR + R)): elimination
//     inside each partition expresses its rows through the last unknown of
//     the previous partition and its own last unknown, the resulting
//     tridiagonal system in those 64 unknowns is solved by parallel cyclic
//     reduction across lanes, and the partition rows follow by substitution.
// The matrix is the one the reference's tridiagonal sweep solves
// (ie.cl:10-48, im.cl:3-41): sub-diagonal a_i = J[i][i+1] for i < N-2 (0
// otherwise), diagonal J[i][i], super-diagonal c_i = J[i+1][i] (c_{N-1} = 0),
// each entry already scaled to I + s*dt*J.  Solving it by partitions instead
// of one sweep changes only the rounding.

#define SODECL_BR ((_numeq_ + SODECL_BG - 1) / SODECL_BG)   // rows per partition == components per lane

// Coefficients of row i of the reference's tridiagonal matrix.
SODECL_INLINE double sodecl_tri_a(const double* up, int i) { return (i >= 1 && i < _numeq_ - 2) ? up[i] : 0.0; }
SODECL_INLINE double sodecl_tri_c(const double* lo, int i) { return (i < _numeq_ - 1) ? lo[i] : 0.0; }

SODECL_INLINE void sodecl_pcr_neighbour(double v, int lane, int s, bool up, double identity, double* out)
{
    const double w = up ? __shfl_up(v, (unsigned)s, SODECL_BG) : __shfl_down(v, (unsigned)s, SODECL_BG);
    const bool inside = up ? lane - s >= 0 : lane + s < SODECL_BG;
    *out = inside ? w : identity;
}

// Solves M x = v for the matrix above; lo, di, up, v, x in LDS; all 64 lanes.
SODECL_INLINE void sodecl_tri_solve(const double* lo, const double* di, const double* up, const double* v, double* x,
                                    int lane)
{
    const int R = SODECL_BR;
    const int r0 = lane * R;
    const bool active = r0 < _numeq_;
    const int r1 = active ? min(_numeq_, r0 + R) - 1 : r0;
    const int nin = active ? r1 - r0 : 0;   // interior rows r0 .. r1-1
    double al[SODECL_BR], be[SODECL_BR], ga[SODECL_BR];
    double e[SODECL_BR], f[SODECL_BR], g[SODECL_BR];
    // forward elimination through the interior, tracking the coupling to
    // x_L = last unknown of the previous partition
#pragma unroll
    for (int j = 0; j < SODECL_BR - 1; ++j) {
        if (j < nin) {
            const int i = r0 + j;
            if (j == 0) {
                al[0] = di[i];
                be[0] = v[i];
                ga[0] = sodecl_tri_a(up, i);
            } else {
                const double m = sodecl_tri_a(up, i) / al[j - 1];
                al[j] = di[i] - m * sodecl_tri_c(lo, i - 1);
                be[j] = v[i] - m * be[j - 1];
                ga[j] = -m * ga[j - 1];
            }
        }
    }
    // backward: interior x_i = e_i + f_i x_L + g_i x_R, x_R = own last unknown
#pragma unroll
    for (int j = SODECL_BR - 2; j >= 0; --j) {
        if (j < nin) {
            const int i = r0 + j;
            const double inv = 1.0 / al[j];
            const double ci = sodecl_tri_c(lo, i);
            if (j == nin - 1) {
                e[j] = be[j] * inv;
                f[j] = -ga[j] * inv;
                g[j] = -ci * inv;
            } else {
                e[j] = (be[j] - ci * e[j + 1]) * inv;
                f[j] = (-ga[j] - ci * f[j + 1]) * inv;
                g[j] = (-ci * g[j + 1]) * inv;
            }
        }
    }
    // reduced row of partition k (row r1): A X_{k-1} + B X_k + C X_{k+1} = D
    const bool has_in = nin > 0;
    const double e_last = has_in ? e[nin - 1] : 0.0;
    const double f_last = has_in ? f[nin - 1] : 1.0;   // x_{r1-1} is X_{k-1} itself without interior
    const double g_last = has_in ? g[nin - 1] : 0.0;
    // the next partition's first row, through X_k and X_{k+1}
    const double e_first = has_in ? e[0] : 0.0;
    const double f_first = has_in ? f[0] : 0.0;
    const double g_first = has_in ? g[0] : 1.0;   // without interior its first row is its X
    const double e_nx = __shfl_down(e_first, 1u, SODECL_BG);
    const double f_nx = __shfl_down(f_first, 1u, SODECL_BG);
    const double g_nx = __shfl_down(g_first, 1u, SODECL_BG);
    double A = 0.0, B = 1.0, C = 0.0, D = 0.0;
    if (active) {
        const double ar = sodecl_tri_a(up, r1);   // zero for row 0
        const double cr = sodecl_tri_c(lo, r1);
        const bool has_next = r1 + 1 < _numeq_;
        A = ar * f_last;
        B = di[r1] + ar * g_last + (has_next ? cr * f_nx : 0.0);
        C = has_next ? cr * g_nx : 0.0;
        D = v[r1] - ar * e_last - (has_next ? cr * e_nx : 0.0);
    }
    // parallel cyclic reduction over the 64 reduced rows
#pragma unroll
    for (int s = 1; s < SODECL_BG; s <<= 1) {
        double Am, Bm, Cm, Dm, Ap, Bp, Cp, Dp;
        sodecl_pcr_neighbour(A, lane, s, true, 0.0, &Am);
        sodecl_pcr_neighbour(B, lane, s, true, 1.0, &Bm);
        sodecl_pcr_neighbour(C, lane, s, true, 0.0, &Cm);
        sodecl_pcr_neighbour(D, lane, s, true, 0.0, &Dm);
        sodecl_pcr_neighbour(A, lane, s, false, 0.0, &Ap);
        sodecl_pcr_neighbour(B, lane, s, false, 1.0, &Bp);
        sodecl_pcr_neighbour(C, lane, s, false, 0.0, &Cp);
        sodecl_pcr_neighbour(D, lane, s, false, 0.0, &Dp);
        const double k1 = A / Bm;
        const double k2 = C / Bp;
        const double nA = -k1 * Am;
        const double nC = -k2 * Cp;
        const double nB = B - k1 * Cm - k2 * Ap;
        const double nD = D - k1 * Dm - k2 * Dp;
        A = nA;
        B = nB;
        C = nC;
        D = nD;
    }
    const double X = D / B;
    double Xm;
    sodecl_pcr_neighbour(X, lane, 1, true, 0.0, &Xm);
    if (active) {
#pragma unroll
        for (int j = 0; j < SODECL_BR - 1; ++j)
            if (j < nin) x[r0 + j] = e[j] + f[j] * Xm + g[j] * X;
        x[r1] = X;
    }
}

// Integrates one orbit for `steps` steps of implicit Euler (SODECL_SOLVER 3,
// ie.cl:50-117) or implicit midpoint (4, im.cl:43-118) with the reference's
// Newton iteration: stop when the L1 (IE) / L2 (IM) residual norm is below
// _epsilon1_, at most SODECL_NEWTON_MAXIT updates.  Snapshot j (state after
// j * _numsteps_ steps) goes to out[j * neq ...].  Returns the Newton
// updates taken.
__device__ long long sodecl_batch_orbit(const double* __restrict__ y0, const double* __restrict__ prm,
                                        double* __restrict__ out, int steps, double* lds)
{
    const int lane = threadIdx.x;
    double* const sh_g = lds;                  // evaluation point (g, or (y+g)/2)
    double* const sh_r = lds + _numeq_;        // residual -> right-hand side of the solve
    double* const sh_lo = lds + 2 * _numeq_;
    double* const sh_di = lds + 3 * _numeq_;
    double* const sh_up = lds + 4 * _numeq_;
    double* const sh_x = lds + 5 * _numeq_;
    double* const sh_p = lds + 6 * _numeq_;
    for (int k = lane; k < _numpar_; k += SODECL_BG) sh_p[k] = prm[k];
    double y[SODECL_BR], gg[SODECL_BR];
#pragma unroll
    for (int c = 0; c < SODECL_BR; ++c) {
        const int comp = lane + c * SODECL_BG;
        y[c] = comp < _numeq_ ? y0[comp] : 0.0;
        if (comp < _numeq_) out[comp] = y[c];
    }
    double t = 0.0;
    long long updates = 0;
#if SODECL_SOLVER == 3
    const double scale = -1.00;
#else
    const double scale = -0.50;
#endif
    for (int s = 0; s < steps; ++s) {
#pragma unroll
        for (int c = 0; c < SODECL_BR; ++c) gg[c] = y[c];
        for (int it = 0; it < SODECL_NEWTON_MAXIT; ++it) {
            __syncthreads();   // previous readers of sh_g are done
#pragma unroll
            for (int c = 0; c < SODECL_BR; ++c) {
                const int comp = lane + c * SODECL_BG;
#if SODECL_SOLVER == 3
                if (comp < _numeq_) sh_g[comp] = gg[c];
#else
                if (comp < _numeq_) sh_g[comp] = (y[c] + gg[c]) * 0.5;
#endif
            }
            __syncthreads();
            double r[SODECL_BR];
            double err = 0.0;
#pragma unroll
            for (int c = 0; c < SODECL_BR; ++c) {
                const int comp = lane + c * SODECL_BG;
                r[c] = 0.0;
                if (comp < _numeq_) {
                    const double fv = sode_system_i(comp, t, sh_g, sh_p);
                    r[c] = gg[c] - y[c] - fv * _m_dt_;
#if SODECL_SOLVER == 3
                    err = err + fabs(r[c]);
#else
                    err = err + r[c] * r[c];
#endif
                }
            }
            err = sodecl_group_sum(err);
#if SODECL_SOLVER == 3
            if (err < _epsilon1_) break;
#else
            if (sqrt(err) < _epsilon1_) break;
#endif
#pragma unroll
            for (int c = 0; c < SODECL_BR; ++c) {
                const int comp = lane + c * SODECL_BG;
                if (comp < _numeq_) {
                    double lo, di, up;
                    calc_jacobian_band_i(comp, sh_g, sh_p, &lo, &di, &up);
                    sh_lo[comp] = lo * scale * _m_dt_;
                    sh_up[comp] = up * scale * _m_dt_;
                    sh_di[comp] = di * scale * _m_dt_ + 1;
                    sh_r[comp] = r[c];
                }
            }
            __syncthreads();
            sodecl_tri_solve(sh_lo, sh_di, sh_up, sh_r, sh_x, lane);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < SODECL_BR; ++c) {
                const int comp = lane + c * SODECL_BG;
                if (comp < _numeq_) gg[c] = gg[c] - sh_x[comp];
            }
            ++updates;
        }
#pragma unroll
        for (int c = 0; c < SODECL_BR; ++c) y[c] = gg[c];
        t = t + _m_dt_;
        if ((s + 1) % _numsteps_ == 0) {
            double* o = out + (size_t)((s + 1) / _numsteps_) * _numeq_;
#pragma unroll
            for (int c = 0; c < SODECL_BR; ++c) {
                const int comp = lane + c * SODECL_BG;
                if (comp < _numeq_) o[comp] = y[c];
            }
        }
    }
    return updates;
}

#undef SODECL_BR
