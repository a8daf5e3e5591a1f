/*
 * Oculomotor model as a plain plugin: the model of the reference's example
 * (examples/oculomotor.cl:1-36, sode_system :83-97) written the way a user
 * writes a system source -- a branch per burst-curve side, IEEE divisions,
 * no shared work between the two burst curves, no device-specific helpers.
 * Each right-hand side performs the same floating-point operations in the
 * same order as the reference's, so its results match the tuned
 * oculomotor.cl plugin to rounding (the tuned plugin evaluates one
 * exponential per side through a selected exponent and divides through
 * reciprocals with a correction step; same bits).
 *
 * It exists to measure what an unmodified user source gets from the
 * integrators and the JIT (the own sine and exponential apply to it too).
 */

#define TAU1 0.15
#define TAU2 0.012
#define TAUN 25.00

double burst_rate(double x, double a, double b, double at, double bt)
{
    if (x > 0)
        return at * (1.00 - exp(-x / bt));
    return -(a / b) * x * exp(x / b);
}

void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    const double rate_sum = (1.00 / TAU1) + (1.00 / TAU2);
    const double rate_prod = 1.00 / (TAU1 * TAU2);
    yout[0] = y[1];
    yout[1] = -(rate_sum * y[1]) - (rate_prod * y[0]) + (rate_prod * y[2]) + rate_sum * (y[3] - y[4]);
    yout[2] = -(y[2] / TAUN) + (y[3] - y[4]);
    yout[3] = (1.00 / p[2]) * (-y[3] - (p[3] * y[3] * y[4] * y[4]) + burst_rate(y[5], p[0], p[1], p[4], p[5]));
    yout[4] = (1.00 / p[2]) * (-y[4] - (p[3] * y[4] * y[3] * y[3]) + burst_rate(-y[5], p[0], p[1], p[4], p[5]));
    yout[5] = -(y[3] - y[4]);
}
