/*
 * 1-D Brusselator reaction-diffusion chain as a SODECL system plugin.
 *
 * n = _numeq_/2 cells, state interleaved per cell: y[2k] = u_k, y[2k+1] = v_k.
 *
 *   du_k/dt = A - (B+1) u_k + u_k^2 v_k + a (u_{k-1} - 2 u_k + u_{k+1})
 *   dv_k/dt = B u_k - u_k^2 v_k         + a (v_{k-1} - 2 v_k + v_{k+1})
 *
 * Unit cell spacing; Dirichlet ends held at the homogeneous steady state
 * (u = A, v = B/A) through ghost cells u_{-1} = u_n = A, v_{-1} = v_n = B/A.
 * Parameters per orbit: p0 = A, p1 = B, p2 = a (diffusion).
 *
 * Entry points: sode_system (drift), calc_jacobian (dense, as the implicit
 * solvers of the reference read it), calc_jacobian_band (the three diagonals
 * the tridiagonal solve reads: lower[i] = J[i+1][i], diag[i] = J[i][i],
 * upper[i] = J[i][i+1]) and sode_system_i (one component).
 */

#define BRU_CELLS (_numeq_ / 2)

/* Component i: reaction term of its species plus diffusion along the chain
 * of the same species (neighbours i - 2 and i + 2, ghost values at the
 * ends).  Written without a branch on the species, so the lanes of a
 * wavefront that hold u and v components follow one instruction stream;
 * only the end cells take the ghost-value branch. */
double sode_system_i(int i, double t, double y[_numeq_], double p[_numpar_])
{
    const double A = p[0];
    const double B = p[1];
    const double a = p[2];
    const int k = i / 2;
    const int odd = i & 1;
    const double u = y[2 * k];
    const double v = y[2 * k + 1];
    const double uuv = u * u * v;
    const double self = odd ? v : u;
    double left, right;
    if (k > 0 && k < BRU_CELLS - 1)
    {
        left = y[i - 2];
        right = y[i + 2];
    }
    else
    {
        const double ghost = odd ? B / A : A;
        left = (k > 0) ? y[i - 2] : ghost;
        right = (k < BRU_CELLS - 1) ? y[i + 2] : ghost;
    }
    const double react = odd ? B * u - uuv : A - (B + 1) * u + uuv;
    return react + a * (left - 2 * self + right);
}

void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
    {
        yout[i] = sode_system_i(i, t, y, p);
    }
}

/* Row i of the three diagonals; the off-band diffusion couplings at
 * distance 2 are not part of the band. */
void calc_jacobian_band_i(int i, double y[_numeq_], double p[_numpar_], double* lower, double* diag, double* upper)
{
    const double B = p[1];
    const double a = p[2];
    const int k = i / 2;
    const int odd = i & 1;
    const double u = y[2 * k];
    const double v = y[2 * k + 1];
    /* even i (u_k):  J[2k][2k] = -(B+1) + 2uv - 2a, J[2k][2k+1] = d(du_k)/dv_k = u^2,
     *                J[2k+1][2k] = d(dv_k)/du_k = B - 2uv;
     * odd i (v_k):   J[2k+1][2k+1] = -u^2 - 2a, and no coupling to the next
     *                cell's u (J[2k+1][2k+2] = J[2k+2][2k+1] = 0) */
    *diag = odd ? -(u * u) - 2 * a : -(B + 1) + 2 * u * v - 2 * a;
    *upper = odd ? 0.0 : u * u;
    *lower = odd ? 0.0 : B - 2 * u * v;
}

void calc_jacobian_band(double y[_numeq_], double lower[_numeq_], double diag[_numeq_], double upper[_numeq_], double p[_numpar_])
{
    for (int i = 0; i < _numeq_; i++)
    {
        calc_jacobian_band_i(i, y, p, &lower[i], &diag[i], &upper[i]);
    }
    lower[_numeq_ - 1] = 0;
    upper[_numeq_ - 1] = 0;
}

void calc_jacobian(double y[_numeq_], double jac[_numeq_][_numeq_], double p[_numpar_])
{
    const double a = p[2];
    for (int i = 0; i < _numeq_; i++)
    {
        for (int j = 0; j < _numeq_; j++)
        {
            jac[i][j] = 0;
        }
    }
    for (int i = 0; i < _numeq_; i++)
    {
        double lo, di, up;
        calc_jacobian_band_i(i, y, p, &lo, &di, &up);
        jac[i][i] = di;
        if (i + 1 < _numeq_)
        {
            jac[i][i + 1] = up;
            jac[i + 1][i] = lo;
        }
        /* diffusion to the same species in the neighbouring cells */
        if (i - 2 >= 0)
        {
            jac[i][i - 2] = a;
        }
        if (i + 2 < _numeq_)
        {
            jac[i][i + 2] = a;
        }
    }
}
