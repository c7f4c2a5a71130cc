// euler_device.hpp -- pointwise device routines of the 2-D Euler system and
// the Fortran intrinsic semantics the kernels must reproduce.
//
// Association order follows the reference expressions exactly so that the
// build without floating-point contraction is bit-identical to the reference
// routines (checked by the GPU parity tests).
#pragma once
#include <hip/hip_runtime.h>

#define ADG_DEV __device__ __forceinline__

namespace aderdg {

constexpr int kV = 4;  // nVar of the Euler system (Euler/ader_dg.f90:20)

// Two-argument MAX/MIN as flang evaluates them: (a > b ? a : b), so a NaN
// second argument wins and a NaN first argument loses.
ADG_DEV double f_max(double a, double b) { return a > b ? a : b; }
ADG_DEV double f_min(double a, double b) { return a < b ? a : b; }
// One element of MAXVAL/MINVAL, which skip NaN elements; start from NaN.
ADG_DEV double fold_max(double m, double e)
{
    return __builtin_isnan(e) ? m : ((__builtin_isnan(m) || e > m) ? e : m);
}
ADG_DEV double fold_min(double m, double e)
{
    return __builtin_isnan(e) ? m : ((__builtin_isnan(m) || e < m) ? e : m);
}

// PDEFlux, Euler/pde.f90:46-69.  F[0..3] x-flux, F[4..7] y-flux.
ADG_DEV void euler_flux(double g, const double *Q, double *F)
{
    const double irho = 1.0 / Q[0];
    const double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
    F[0] = Q[1];
    F[1] = irho * Q[1] * Q[1] + p;
    F[2] = irho * Q[1] * Q[2];
    F[3] = irho * Q[1] * (Q[3] + p);
    F[4] = Q[2];
    F[5] = irho * Q[2] * Q[1];
    F[6] = irho * Q[2] * Q[2] + p;
    F[7] = irho * Q[2] * (Q[3] + p);
}

// maxval(abs(Lambda)) of PDEEigenvalues, Euler/pde.f90:75-91, normal (nx, ny)
ADG_DEV double euler_maxabs_eig(double g, const double *Q, double nx, double ny)
{
    const double u = (Q[1] * nx + Q[2] * ny) / Q[0];
    const double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]);
    const double c = __builtin_sqrt(g * p / Q[0]);
    double m = __builtin_nan("");
    m = fold_max(m, __builtin_fabs(u - c));
    m = fold_max(m, __builtin_fabs(u));
    m = fold_max(m, __builtin_fabs(u));
    m = fold_max(m, __builtin_fabs(u + c));
    return m;
}

// RusanovFlux, Euler/pde.f90:97-114 (face fluxes given)
ADG_DEV void euler_rusanov(double g, const double *QL, const double *FL, const double *QR, const double *FR,
                           double nx, double ny, double *out)
{
    const double smax = f_max(euler_maxabs_eig(g, QL, nx, ny), euler_maxabs_eig(g, QR, nx, ny));
#pragma unroll
    for (int v = 0; v < kV; v++) out[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * smax * (QR[v] - QL[v]);
}

// RusanovFluxS, Euler/limiter.f90:361-378 (fluxes computed from the states)
ADG_DEV void euler_rusanov_s(double g, const double *QL, const double *QR, double nx, double ny, double *out)
{
    const double smax = f_max(euler_maxabs_eig(g, QL, nx, ny), euler_maxabs_eig(g, QR, nx, ny));
    double FL[2 * kV], FR[2 * kV];
    euler_flux(g, QL, FL);
    euler_flux(g, QR, FR);
#pragma unroll
    for (int v = 0; v < kV; v++) {
        double m = 0.0;
        m = m + (FL[v] + FR[v]) * nx;
        m = m + (FL[kV + v] + FR[kV + v]) * ny;
        out[v] = 0.5 * m - 0.5 * smax * (QR[v] - QL[v]);
    }
}

// PDEAssurePositivity, Euler/pde.f90:120-145: true if the state is rejected
ADG_DEV bool euler_inadmissible(double g, const double *Q)
{
    const double irho = 1.0 / Q[0];
    const double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
    return (Q[0] <= 1.0e-12) || (p <= 1.0e-12);
}

// minmod, Euler/limiter.f90:384-405
ADG_DEV double minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    return __builtin_fabs(a) < __builtin_fabs(b) ? a : b;
}

// Neighbour along one axis: clamp to the mesh, or wrap when periodic
// (Euler/initialize.f90:231-255).
ADG_DEV int nb_index(int i, int d, int n, int wrap)
{
    int a = i + d;
    if (a < 0) a = wrap ? n - 1 : 0;
    if (a > n - 1) a = wrap ? 0 : n - 1;
    return a;
}

}  // namespace aderdg
