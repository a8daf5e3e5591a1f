/*
 * Oculomotor (saccadic eye movement) model as a SODECL system plugin.
 *
 * States (6):  y0 = g (eye position), y1 = u (eye velocity), y2 = n
 *              (neural integrator), y3 = r and y4 = l (right/left burst
 *              neurons), y5 = m (motor error).
 * Parameters (6 per orbit):
 *              p0, p1  burst-curve amplitude and width for negative input,
 *              p2      burst-neuron time constant,
 *              p3      cross-inhibition strength between r and l,
 *              p4, p5  burst-curve amplitude and width for positive input.
 * SDE variant: two extra parameters p[_numpar_-2], p[_numpar_-1] scale the
 *              two noise processes entering the velocity equation.
 *
 * The model equations are those of examples/oculomotor.cl:1-112 in the
 * reference (sode_system :83-97, sode_system_stoch :99-112, calc_jacobian
 * :38-81, burst curve and its derivative :5-36).  Each right-hand side is
 * written with the same floating-point operations in the same order, so
 * trajectories agree with the reference to rounding.
 *
 * Plugin contract: OpenCL-C style functions over fixed-size arrays whose
 * extents are the macros _numeq_, _numpar_ and _numnoi_ supplied by the
 * integrator.  Optional entry points: calc_jacobian_band, read by the
 * implicit solvers, returns only the three diagonals the tridiagonal solve
 * uses (lower[i] = J[i+1][i], diag[i] = J[i][i], upper[i] = J[i][i+1]);
 * sode_system_team, read by the team schedule, shares the two burst curves
 * between the two lanes of a pair; calc_jacobian_exact and
 * calc_jacobian_params (df/dy and df/dp) are read by the discrete adjoint of
 * the implicit solvers.
 */

#define OCU_T1 0.15
#define OCU_T2 0.012
#define OCU_TN 25.00

/* a / d given rd = 1/d correctly rounded (d a constant or an orbit
 * parameter, so rd is computed once): q0 = a*rd, the exact residual
 * q0*d - a by one fused multiply-add, one correction (Markstein) -- the bits
 * of IEEE a / d outside the subnormal range, without a general division per
 * call.  No select for special values: with the residual's sign taken this
 * way a zero quotient keeps its sign; an infinite dividend or an overflowing
 * q0 gives a NaN or -inf residual, which the clamp to -DBL_MAX makes finite so
 * that the correction leaves q0 in place; NaN stays NaN.  That holds for
 * d >= 1 or a dividend whose sign bit is set (every call here: the constant
 * Tn and the burst curve's -|x|). */
double ocu_div(double a, double d, double rd)
{
#if defined(SODECL_STRICT) && !SODECL_STRICT
    /* the GPU's default arithmetic (contraction on): the product with the
     * reciprocal, within an ulp of the quotient -- the size of rounding
     * change contraction makes anyway; right for infinities, NaN and
     * signed zeros.  Strict arithmetic, and every host build, keep the
     * quotient's bits. */
    return a * rd;
#else
    const double q0 = a * rd;
    const double r = fmax(fma(q0, d, -a), -0x1.fffffffffffffp+1023);
    return fma(-r, rd, q0);
#endif
}

/* Burst-neuron firing curve: saturating exponential for x > 0, a scaled
 * x*exp(x/b) ramp otherwise.  The exponent is selected before the single
 * exp() call, so lanes of a wavefront that take different sides still
 * evaluate one exponential; each side computes exactly the operations of
 * the two-branch form ((-x)/bt or x/b, then exp); numerator and denominator
 * are selected first so the division is evaluated once as well. */
double ocu_burst(double x, double a, double b, double at, double bt)
{
    const int pos = x > 0;
    /* one division through the selected parameter's reciprocal; both
     * reciprocals depend on the orbit only and are computed once per run.
     * (2/bt)*0.5 is exactly 1/bt (scaling by powers of two commutes with
     * rounding); written so, the two candidates are not the same operation
     * on different operands, which a compiler would turn into one division
     * of the selected denominator inside the step loop. */
    const double rbt = (2.0 / bt) * 0.5;
    const double rb = 1.0 / b;
    /* the numerator -x (x > 0) or x (x <= 0) is -|x|, a free source
     * modifier; for x = +0 it is -0, whose quotient's exp is 1 as well */
    const double e = exp(ocu_div(-fabs(x), pos ? bt : b, pos ? rbt : rb));
    return pos ? at * (1.00 - e) : -(a / b) * x * e;
}

/* Both burst curves of the right-hand side, B(x) and B(-x), with their
 * shared work done once.  For x > 0, B(x) is on its saturating side with
 * exponent (-x)/bt = -|x|/bt and B(-x) on its ramp side with exponent
 * (-x)/b = -|x|/b; for x < 0 the two swap; for a zero x both are ramps whose
 * exponent is a zero, whose exp is 1 like exp(-|x|/b).  Every case therefore
 * needs exactly exp(-|x|/bt) and exp(-|x|/b).  The saturating value
 * at(1 - exp(-|x|/bt)) serves whichever curve is on that side, and the ramp
 * values -(a/b) x exp(-|x|/b) of B(x) and -(a/b) (-x) exp(-|x|/b) of B(-x)
 * are negatives of each other bit for bit (rounding to nearest is symmetric
 * in sign).  Each result has the bits of ocu_burst's. */
void ocu_burst_pair(double x, double a, double b, double at, double bt, double* bx, double* bmx)
{
    const double sat = at * (1.00 - exp(ocu_div(-fabs(x), bt, 1.0 / bt)));
    const double ramp = -(a / b) * x * exp(ocu_div(-fabs(x), b, 1.0 / b));
    *bx = x > 0 ? sat : ramp;
    *bmx = x < 0 ? sat : -ramp;
}

/* d(ocu_burst)/dx */
double ocu_burst_dx(double x, double a, double b, double at, double bt)
{
    const int pos = x > 0;
    const double e = exp((pos ? -x : x) / (pos ? bt : b));
    return pos ? (at * e) / bt : -((a * e) / b + (a * x * e) / (b * b));
}

/* The right-hand side given the two burst-curve values br = B(y5) and
 * bl = B(-y5). */
void ocu_rhs(double y[_numeq_], double yout[_numeq_], double p[_numpar_], double br, double bl)
{
    const double rate_sum = (1.00 / OCU_T1) + (1.00 / OCU_T2);
    const double rate_prod = 1.00 / (OCU_T1 * OCU_T2);
    const double drive = y[3] - y[4];

    yout[0] = y[1];
    yout[1] = -(rate_sum * y[1]) - (rate_prod * y[0]) + (rate_prod * y[2]) + rate_sum * drive;
    yout[2] = -ocu_div(y[2], OCU_TN, 1.0 / OCU_TN) + drive;
    yout[3] = (1.00 / p[2]) * (-y[3] - (p[3] * y[3] * y[4] * y[4]) + br);
    yout[4] = (1.00 / p[2]) * (-y[4] - (p[3] * y[4] * y[3] * y[3]) + bl);
    yout[5] = -drive;
}

void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    double br, bl;
    ocu_burst_pair(y[5], p[0], p[1], p[4], p[5], &br, &bl);
    ocu_rhs(y, yout, p, br, bl);
}

#ifdef SODECL_TEAM_API
/* Team form of sode_system for the team schedule: the two burst curves, each
 * an exponential and most of the cost of the right-hand side, go to the two
 * lanes of a pair (even lane B(y5), odd lane B(-y5)) and are swapped between
 * them; every lane then forms the whole right-hand side.  The operations are
 * those of sode_system. */
void sode_system_team(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    const int odd = sodecl_lane() & 1;
    const double mine = ocu_burst(odd ? -y[5] : y[5], p[0], p[1], p[4], p[5]);
    const double other = sodecl_group_xor(mine, 1);
    ocu_rhs(y, yout, p, odd ? other : mine, odd ? mine : other);
}
#endif

void sode_system_stoch(double t, double y[_numeq_], double stoch[_numeq_], double p[_numpar_], double noise[_numnoi_])
{
    const double rate_sum = (1.00 / OCU_T1) + (1.00 / OCU_T2);
    for (int i = 0; i < _numeq_; i++)
    {
        stoch[i] = 0;
    }
    stoch[1] = rate_sum * (y[3] - y[4]) * p[_numpar_ - _numnoi_ + 0] * noise[0] + p[_numpar_ - _numnoi_ + 1] * noise[1];
}

void calc_jacobian(double y[_numeq_], double jac[_numeq_][_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
    {
        for (int j = 0; j < _numeq_; j++)
        {
            jac[i][j] = 0;
        }
    }
    jac[0][1] = 1;

    jac[1][0] = -1 / (OCU_T1 * OCU_T2);
    jac[1][1] = -(1 / OCU_T1) - (1 / OCU_T2);
    jac[1][2] = 1 / (OCU_T1 * OCU_T2);
    jac[1][3] = (1 / OCU_T1) + (1 / OCU_T2);
    jac[1][4] = -(1 / OCU_T1) - (1 / OCU_T2);

    jac[2][2] = -1 / OCU_TN;
    jac[2][3] = 1;
    jac[2][4] = -1;

    jac[3][3] = -(p[3] * (y[4] * y[4]) + 1) / p[2];
    jac[3][4] = -(2 * p[3] * y[4] * y[3]) / p[2];
    jac[3][5] = ocu_burst_dx(y[5], p[0], p[1], p[4], p[5]) / p[2];

    jac[4][3] = -(2 * p[3] * y[4] * y[3]) / p[2];
    jac[4][4] = -(p[3] * (y[3] * y[3]) + 1) / p[2];
    jac[4][5] = ocu_burst_dx(-y[5], p[0], p[1], p[4], p[5]) / p[2];

    jac[5][3] = -1;
    jac[5][4] = 1;
}

void calc_jacobian_band(double y[_numeq_], double lower[_numeq_], double diag[_numeq_], double upper[_numeq_], double p[_numpar_])
{
    /* diag[i] = J[i][i] */
    diag[0] = 0;
    diag[1] = -(1 / OCU_T1) - (1 / OCU_T2);
    diag[2] = -1 / OCU_TN;
    diag[3] = -(p[3] * (y[4] * y[4]) + 1) / p[2];
    diag[4] = -(p[3] * (y[3] * y[3]) + 1) / p[2];
    diag[5] = 0;
    /* upper[i] = J[i][i+1] */
    upper[0] = 1;
    upper[1] = 1 / (OCU_T1 * OCU_T2);
    upper[2] = 1;
    upper[3] = -(2 * p[3] * y[4] * y[3]) / p[2];
    upper[4] = ocu_burst_dx(-y[5], p[0], p[1], p[4], p[5]) / p[2];
    upper[5] = 0;
    /* lower[i] = J[i+1][i] */
    lower[0] = -1 / (OCU_T1 * OCU_T2);
    lower[1] = 0;
    lower[2] = 0;
    lower[3] = -(2 * p[3] * y[4] * y[3]) / p[2];
    lower[4] = 1;
    lower[5] = 0;
}

/* The true df/dy, read by the discrete adjoint of the implicit solvers.  It
 * is calc_jacobian except J[4][5]: yout[4] depends on y5 through B(-y5), so
 * dyout[4]/dy5 = -B'(-y5)/p2.  calc_jacobian keeps the reference's entry
 * (examples/oculomotor.cl:75, without the minus sign) because the Newton
 * iteration of the implicit solvers uses it as its matrix. */
void calc_jacobian_exact(double y[_numeq_], double jac[_numeq_][_numeq_], double p[_numpar_])
{
    calc_jacobian(y, jac, p);
    jac[4][5] = -jac[4][5];
}

/* Partial derivatives of the burst curve B(x; a, b, at, bt) with respect to
 * (a, b, at, bt). */
void ocu_burst_dp(double x, double a, double b, double at, double bt, double d[4])
{
    if (x > 0)
    {
        const double e = exp(-x / bt);
        d[0] = 0;
        d[1] = 0;
        d[2] = 1.00 - e;
        d[3] = -(at * e * x) / (bt * bt);
    }
    else
    {
        const double e = exp(x / b);
        d[0] = -(x / b) * e;
        d[1] = ((a * x * e) / (b * b)) * (1.00 + x / b);
        d[2] = 0;
        d[3] = 0;
    }
}

/* df/dp (jp[i][k] = dyout[i]/dp[k]): only yout[3] and yout[4] depend on the
 * parameters; the noise amplitudes of the SDE form do not enter the drift. */
void calc_jacobian_params(double y[_numeq_], double jp[_numeq_][_numpar_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
    {
        for (int k = 0; k < _numpar_; k++)
        {
            jp[i][k] = 0;
        }
    }
    double d3[4], d4[4];
    ocu_burst_dp(y[5], p[0], p[1], p[4], p[5], d3);
    ocu_burst_dp(-y[5], p[0], p[1], p[4], p[5], d4);
    const double s3 = -y[3] - (p[3] * y[3] * y[4] * y[4]) + ocu_burst(y[5], p[0], p[1], p[4], p[5]);
    const double s4 = -y[4] - (p[3] * y[4] * y[3] * y[3]) + ocu_burst(-y[5], p[0], p[1], p[4], p[5]);
    jp[3][0] = d3[0] / p[2];
    jp[3][1] = d3[1] / p[2];
    jp[3][2] = -s3 / (p[2] * p[2]);
    jp[3][3] = -(y[3] * y[4] * y[4]) / p[2];
    jp[3][4] = d3[2] / p[2];
    jp[3][5] = d3[3] / p[2];
    jp[4][0] = d4[0] / p[2];
    jp[4][1] = d4[1] / p[2];
    jp[4][2] = -s4 / (p[2] * p[2]);
    jp[4][3] = -(y[4] * y[3] * y[3]) / p[2];
    jp[4][4] = d4[2] / p[2];
    jp[4][5] = d4[3] / p[2];
}
