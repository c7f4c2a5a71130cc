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
#include "ic_spec.hpp"
#include "operators.hpp"

#pragma clang fp contract(off)

namespace aderdg {

// The reference catalogue entries as IcSpecs.
bool ic_of_ictype(int ictype, const double *xl, const double *xr, IcSpec *ic)
{
    *ic = IcSpec{};
    if (ictype == 11) {
        const double pi = 4.0 * std::atan(1.0);
        ic->kind = kIcDmr;
        const double p[10] = {1.0 / 6.0, std::tan(pi / 3.0), 8.0, 8.25 * std::cos(pi / 6.0), -8.25 * std::sin(pi / 6.0),
                              116.5, 1.4, 0.0, 0.0, 1.0};
        for (int k = 0; k < 10; k++) ic->p[k] = p[k];
    } else if (ictype == 4) {
        ic->kind = kIcVortex;
        const double p[5] = {5.0, 0.5 * (xl[0] + xr[0]), 0.5 * (xl[1] + xr[1]), 1.0, 1.0};
        for (int k = 0; k < 5; k++) ic->p[k] = p[k];
    } else if (ictype == 5) {  // Lax-Liu configuration 3
        ic->kind = kIcQuadrant;
        const double p[18] = {0.5 * (xl[0] + xr[0]), 0.5 * (xl[1] + xr[1]), 1.5, 0.0, 0.0, 1.5, 0.5323, 1.206, 0.0, 0.3,
                              0.138, 1.206, 1.206, 0.029, 0.5323, 0.0, 1.206, 0.3};
        for (int k = 0; k < 18; k++) ic->p[k] = p[k];
    } else {
        return false;
    }
    return true;
}

// The Baer-Nunziato test problems of Baer-Nunziato/initialize.f90:314-403 (ICType 1, 2,
// 3 and 7): one-dimensional Riemann problems, the left state where x <= 0.  Primitive
// states (rho_s, u_s, v_s, p_s, rho_g, u_g, v_g, p_g, phi_s).
bool bn_ic_of_ictype(int ictype, IcSpec *ic)
{
    static const double states[4][2][9] = {
        {{1.0, 0.0, 0.0, 1.0, 0.5, 0.0, 0.0, 1.0, 0.4}, {2.0, 0.0, 0.0, 2.0, 1.5, 0.0, 0.0, 2.0, 0.8}},
        {{800.0, 0.0, 0.0, 500.0, 1.5, 0.0, 0.0, 2.0, 0.4}, {1000.0, 0.0, 0.0, 600.0, 1.0, 0.0, 0.0, 1.0, 0.3}},
        {{1.0, 0.9, 0.0, 2.5, 1.0, 0.0, 0.0, 1.0, 0.9}, {1.0, 0.0, 0.0, 1.0, 1.2, 1.0, 0.0, 2.0, 0.2}},
        {{800.0, 1.0, 0.0, 1.0, 2.0, 1.0, 0.0, 1.0, 0.99}, {1000.0, 1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 1.0, 0.01}}};
    const int k = ictype == 1 ? 0 : ictype == 2 ? 1 : ictype == 3 ? 2 : ictype == 7 ? 3 : -1;
    *ic = IcSpec{};
    if (k < 0) return false;
    ic->kind = kIcBnRiemann;
    ic->p[0] = 0.0;
    for (int v = 0; v < 9; v++) {
        ic->p[1 + v] = states[k][0][v];
        ic->p[10 + v] = states[k][1][v];
    }
    return true;
}

// uh(V,n,n,IMAX,JMAX) of an IcSpec on the host.
int ic_fill(const IcSpec &ic, int N, int V, int imax, int jmax, const double *xl, const double *xr, double gamma,
            const double *eos, double *uh)
{
    Operators op;
    if (!build_operators(N, op)) return ADERDG_ESINGULAR;
    const int n = N + 1;
    const double dx = (xr[0] - xl[0]) / imax, dy = (xr[1] - xl[1]) / jmax;
    for (int j = 0; j < jmax; j++)
        for (int i = 0; i < imax; i++) {
            const double cx = (xl[0] + ((double)i + 0.5) * dx) - 0.5 * dx;
            const double cy = (xl[1] + ((double)j + 0.5) * dy) - 0.5 * dy;
            double *u = uh + (size_t)V * n * n * ((size_t)i + (size_t)imax * j);
            for (int l = 0; l < n; l++)
                for (int k = 0; k < n; k++) ic_field(ic, gamma, eos, cx + op.xi[k] * dx, cy + op.xi[l] * dy, u + V * (k + n * l));
        }
    return 0;
}

}  // namespace aderdg

extern "C" int aderdg_initial_condition(int ictype, int N, int imax, int jmax, const double *xl, const double *xr,
                                        double gamma, double *uh)
{
    if (N < 1 || N > aderdg::kMaxDegree || imax < 1 || jmax < 1 || !xl || !xr || !uh) return ADERDG_EINVAL;
    aderdg::IcSpec ic;
    if (!aderdg::ic_of_ictype(ictype, xl, xr, &ic)) return ADERDG_EINVAL;
    return aderdg::ic_fill(ic, N, 4, imax, jmax, xl, xr, gamma, nullptr, uh);
}

extern "C" int aderdg_ic_of_ictype(int pde, int ictype, const double *xl, const double *xr, aderdg_ic *out)
{
    if (!xl || !xr || !out) return ADERDG_EINVAL;
    aderdg::IcSpec s{};
    const bool ok = pde == ADERDG_PDE_EULER ? aderdg::ic_of_ictype(ictype, xl, xr, &s)
                    : pde == ADERDG_PDE_BN  ? aderdg::bn_ic_of_ictype(ictype, &s)
                                            : false;
    if (!ok) return ADERDG_EINVAL;
    out->kind = s.kind;
    out->pad = 0;
    for (int k = 0; k < aderdg::kIcParams; k++) out->p[k] = s.p[k];
    return 0;
}

extern "C" int aderdg_ic_eval(const aderdg_desc *d, const aderdg_ic *ic, double *uh)
{
    if (!d || !ic || !uh || d->N < 1 || d->N > aderdg::kMaxDegree || d->imax < 1 || d->jmax < 1) return ADERDG_EINVAL;
    if (!aderdg::ic_fits(d->pde, ic->kind)) return ADERDG_EINVAL;
    aderdg::IcSpec s{};
    s.kind = ic->kind;
    for (int k = 0; k < aderdg::kIcParams; k++) s.p[k] = ic->p[k];
    const int V = d->pde == ADERDG_PDE_BN ? 9 : (d->pde == ADERDG_PDE_AD ? 1 : 4);
    return aderdg::ic_fill(s, d->N, V, d->imax, d->jmax, d->xl, d->xr, d->gamma, d->bn_eos, uh);
}
