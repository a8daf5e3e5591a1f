/*
 * Kuramoto network of _numeq_ phase oscillators as a SODECL system plugin.
 *
 *   dtheta_i/dt = w_i + (K/N) * sum_j sin(theta_j - theta_i)
 *
 * Parameters per orbit: p0 = K (coupling), p1..pN = w_i (natural
 * frequencies); for the SDE form a further _numnoi_ = N amplitudes
 * p[_numpar_ - _numnoi_ + i] scale additive noise on each phase.
 *
 * The right-hand side is the one of examples/kuramoto.cl:3-24 in the
 * reference, with the sum over j taken in the same sequential order and the
 * same scaling (K/N) applied after the sum, so trajectories agree with the
 * reference to rounding.
 *
 * Optional entry points read by the MI355X integrators:
 *   sode_system_i        one component of the right-hand side (lets several
 *                        lanes share one orbit, lane i computing component i);
 *   sode_pair,
 *   sode_system_pairsum  the same drift in pairwise form: the pair term
 *                        sin(y_j - y_i) is exactly antisymmetric, so the
 *                        integrator evaluates each pair once;
 *   sode_system_stoch_i  one component of the diffusion term;
 *   calc_jacobian        dense df/dy (implicit solvers, sensitivity checks);
 *   calc_jacobian_params dense df/dp (sensitivity checks).
 */

#define KURA_N _numeq_

double sode_system_i(int i, double t, double y[_numeq_], double p[_numpar_])
{
    double acc = 0;
    for (int j = 0; j < KURA_N; j++)
    {
        acc = acc + sin(y[j] - y[i]);
    }
    acc = acc * (double)(p[0] / (double)(KURA_N));
    return acc + p[i + 1];
}

/* Pairwise form: f_i = (sum over j in order of sode_pair(i, j)) * (K/N) + w_i,
 * the same operations as sode_system_i.  sin is odd and y_j - y_i is the
 * exact negation of y_i - y_j, so sode_pair(i, j) == -sode_pair(j, i); the
 * integrator takes sode_pair(i, i) = sin(0) = +0 without calling it. */
double sode_pair(int i, int j, double t, double y[_numeq_], double p[_numpar_])
{
    return sin(y[j] - y[i]);
}

double sode_system_pairsum(int i, double t, double y[_numeq_], double p[_numpar_], double s)
{
    s = s * (double)(p[0] / (double)(KURA_N));
    return s + p[i + 1];
}

void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < KURA_N; i++)
    {
        yout[i] = sode_system_i(i, t, y, p);
    }
}

double sode_system_stoch_i(int i, double t, double y[_numeq_], double p[_numpar_], double noise[_numnoi_])
{
    return p[_numpar_ - _numnoi_ + i] * noise[i];
}

void sode_system_stoch(double t, double y[_numeq_], double stoch[_numeq_], double p[_numpar_], double noise[_numnoi_])
{
    for (int i = 0; i < KURA_N; i++)
    {
        stoch[i] = sode_system_stoch_i(i, t, y, p, noise);
    }
}

/* J[i][m] = (K/N) * (cos(y_m - y_i) - delta_im * sum_j cos(y_j - y_i)) */
void calc_jacobian(double y[_numeq_], double jac[_numeq_][_numeq_], double p[_numpar_])
{
    const double kn = p[0] / (double)(KURA_N);
    for (int i = 0; i < KURA_N; i++)
    {
        double diag = 0;
        for (int m = 0; m < KURA_N; m++)
        {
            double c = cos(y[m] - y[i]);
            jac[i][m] = kn * c;
            diag = diag + c;
        }
        jac[i][i] = jac[i][i] - kn * diag;
    }
}

/* jp[i][k] = d f_i / d p_k: column 0 is (1/N) sum_j sin(y_j - y_i), column
 * i+1 is 1, the noise amplitudes do not enter the drift. */
void calc_jacobian_params(double y[_numeq_], double jp[_numeq_][_numpar_], double p[_numpar_])
{
    for (int i = 0; i < KURA_N; i++)
    {
        double s = 0;
        for (int j = 0; j < KURA_N; j++)
        {
            s = s + sin(y[j] - y[i]);
        }
        for (int k = 0; k < _numpar_; k++)
        {
            jp[i][k] = 0;
        }
        jp[i][0] = s / (double)(KURA_N);
        jp[i][i + 1] = 1;
    }
}

#ifdef SODECL_GROUP_API
/* Vector-Jacobian product of the drift for the adjoint, evaluated by all
 * lanes of the orbit's group together (x, lam in memory shared by the
 * group).  With c_i = cos x_i, s_i = sin x_i and the group sums
 *   A = sum lam_i c_i, B = sum lam_i s_i, C = sum c_i, D = sum s_i,
 * sin(x_j - x_i) = s_j c_i - c_j s_i gives
 *   (lam^T df/dx)_m = (K/N) (c_m (A - lam_m C) + s_m (B - lam_m D)),
 *   lam^T df/dK     = (A D - B C) / N,      lam^T df/dw_m = lam_m,
 * so each lane does O(N/G) trigonometry and four group reductions instead
 * of the O(N^2) sums of the drift itself. */
void sode_system_vjp_group(double t, double x[_numeq_], double lam[_numeq_], double p[_numpar_], double gx[_numeq_],
                           double gp[_numpar_])
{
    const int G = sodecl_group_size();
    const int l = sodecl_lane();
    double A = 0, B = 0, C = 0, D = 0;
    for (int i = l; i < KURA_N; i += G)
    {
        double s, c;
        sincos(x[i], &s, &c);
        A = A + lam[i] * c;
        B = B + lam[i] * s;
        C = C + c;
        D = D + s;
    }
    A = sodecl_group_sum(A);
    B = sodecl_group_sum(B);
    C = sodecl_group_sum(C);
    D = sodecl_group_sum(D);
    const double kn = p[0] / (double)(KURA_N);
    for (int m = l; m < KURA_N; m += G)
    {
        double s, c;
        sincos(x[m], &s, &c);
        gx[m] = kn * (c * (A - lam[m] * C) + s * (B - lam[m] * D));
        gp[m + 1] = gp[m + 1] + lam[m];
    }
    if (l == 0)
    {
        gp[0] = gp[0] + (A * D - B * C) / (double)(KURA_N);
    }
}
#endif
