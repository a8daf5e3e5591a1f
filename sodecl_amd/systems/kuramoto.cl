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
