// initial.cpp -- initial conditions at the Gauss-Legendre nodes (host).
//
// uh(:,k,l,i,j) = u0(x_GP) with x_GP = corner + xi_k dx, the projection of
// Euler/initialize.f90:262-273 (the nodal basis through the quadrature nodes
// makes the L2 projection a point evaluation).  ICType numbers follow the
// reference catalogue (Euler/params.ini:17-85):
//    4  isentropic vortex: strength 5 on (rho,u,v,p) = (1,1,1,1), centred in
//       the domain  (named at Euler/initialize.f90:58-59, not implemented at HEAD)
//    5  2-D Riemann problem 3 of Lax & Liu, split at the domain centre
//       (named at Euler/initialize.f90:60-61, not implemented at HEAD)
//   11  Double Mach Reflection, InitialField of Euler/initialize.f90:311-341
#include <cmath>

#include "../../include/aderdg.h"
#include "operators.hpp"

#pragma clang fp contract(off)

namespace {

void prim2cons(double g, const double *V, double *Q)  // PDEPrim2Cons, Euler/pde.f90:29-40
{
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = V[3] / (g - 1.0) + 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
}

void field(int ictype, double x, double y, const double *xl, const double *xr, double g, double *Q)
{
    double V[4];
    if (ictype == 11) {
        const double pi = 4.0 * std::atan(1.0);
        if (x <= 1.0 / 6.0 + y / std::tan(pi / 3.0)) {
            V[0] = 8.0;
            V[1] = 8.25 * std::cos(pi / 6.0);
            V[2] = -8.25 * std::sin(pi / 6.0);
            V[3] = 116.5;
        } else {
            V[0] = 1.4;
            V[1] = 0.0;
            V[2] = 0.0;
            V[3] = 1.0;
        }
    } else if (ictype == 4) {
        const double pi = M_PI, beta = 5.0;
        const double xc = 0.5 * (xl[0] + xr[0]), yc = 0.5 * (xl[1] + xr[1]);
        const double dxr = x - xc, dyr = y - yc, r2 = dxr * dxr + dyr * dyr;
        const double dT = -(g - 1.0) * beta * beta / (8.0 * g * pi * pi) * std::exp(1.0 - r2);
        const double rho = std::pow(1.0 + dT, 1.0 / (g - 1.0));
        const double f = beta / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
        V[0] = rho;
        V[1] = 1.0 - f * dyr;
        V[2] = 1.0 + f * dxr;
        V[3] = std::pow(rho, g);
    } else {  // 5: Lax-Liu configuration 3
        const double x0 = 0.5 * (xl[0] + xr[0]), y0 = 0.5 * (xl[1] + xr[1]);
        const bool east = x > x0, north = y > y0;
        static const double NE[4] = {1.5, 0.0, 0.0, 1.5}, NW[4] = {0.5323, 1.206, 0.0, 0.3},
                            SW[4] = {0.138, 1.206, 1.206, 0.029}, SE[4] = {0.5323, 0.0, 1.206, 0.3};
        const double *s = east ? (north ? NE : SE) : (north ? NW : SW);
        for (int v = 0; v < 4; v++) V[v] = s[v];
    }
    prim2cons(g, V, Q);
}

}  // namespace

extern "C" int aderdg_initial_condition(int ictype, int N, int imax, int jmax, const double *xl, const double *xr,
                                        double gamma, double *uh)
{
    if (ictype != 4 && ictype != 5 && ictype != 11) return ADERDG_EINVAL;
    if (N < 1 || N > aderdg::kMaxDegree || imax < 1 || jmax < 1 || !xl || !xr || !uh) return ADERDG_EINVAL;
    aderdg::Operators op;
    if (!aderdg::build_operators(N, op)) return ADERDG_ESINGULAR;
    const int n = N + 1;
    const double dx = (xr[0] - xl[0]) / imax, dy = (xr[1] - xl[1]) / jmax;
    for (int j = 0; j < jmax; j++)
        for (int i = 0; i < imax; i++) {
            const double cx = (xl[0] + ((double)i + 0.5) * dx) - 0.5 * dx;
            const double cy = (xl[1] + ((double)j + 0.5) * dy) - 0.5 * dy;
            double *u = uh + (size_t)4 * n * n * ((size_t)i + (size_t)imax * j);
            for (int l = 0; l < n; l++)
                for (int k = 0; k < n; k++) field(ictype, cx + op.xi[k] * dx, cy + op.xi[l] * dy, xl, xr, gamma, u + 4 * (k + n * l));
        }
    return 0;
}
