// ic_spec.hpp -- initial conditions as point functions shared by the host
// (initial.cpp) and the device (k_init in step_kernels.hip).
//
// The projection of Euler/initialize.f90:262-273 onto the nodal basis through
// the Gauss-Legendre nodes is a point evaluation, so uh(:,k,l,i,j) is the
// field at x = corner_i + xi_k dx, y = corner_j + xi_l dy with the reference's
// node arithmetic: x_c = xL + (i + 0.5) dx (:151), corner = x_c - 0.5 dx (:265).
//
// Parameter layouts (aderdg_ic.p of include/aderdg.h):
//   VORTEX    (Euler) beta, xc, yc, u0, v0: isentropic vortex on rho = p = 1,
//             dT = -(g-1) b^2/(8 g pi^2) e^(1-r^2), rho = (1+dT)^(1/(g-1)),
//             (u,v) = (u0,v0) + b/(2 pi) e^((1-r^2)/2) (-(y-yc), x-xc), p = rho^g
//   QUADRANT  (Euler) x0, y0, then (rho,u,v,p) of NE, NW, SW, SE
//   DMR       (Euler) x_split0, tan(pi/3), then (rho,u,v,p) left and right of
//             x <= x_split0 + y/tan(pi/3) (InitialField, Euler/initialize.f90:311-341)
//   BN_CIRCLE (Baer-Nunziato) xc, yc, radius, p_in, p_out, rho_s, rho_g,
//             phi_in, phi_out: both phases at rest
//   BN_DISC   (Baer-Nunziato) xc, yc, radius, then the primitive states
//             (rho_s,u_s,v_s,p_s,rho_g,u_g,v_g,p_g,phi_s) inside and outside
//   SINE      (advection-diffusion) A, kx, ky: A sin(2 pi kx x) sin(2 pi ky y)
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADG_HD __host__ __device__ inline
#else
#define ADG_HD inline
#endif

namespace aderdg {

enum { kIcVortex = 1, kIcQuadrant = 2, kIcDmr = 3, kIcBnCircle = 4, kIcBnDisc = 5, kIcSine = 6, kIcBnRiemann = 7 };
constexpr int kIcParams = 24;

struct IcSpec {
    int kind;
    int pad;
    double p[kIcParams];
};

// PDEPrim2Cons, Euler/pde.f90:29-40
ADG_HD void ic_euler_cons(double g, const double *V, double *Q)
{
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = V[3] / (g - 1.0) + 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
}

// PDEPrim2Cons, Baer-Nunziato/pde.f90:49-68; V = (rho_s,u_s,v_s,p_s,rho_g,u_g,v_g,p_g,phi_s)
ADG_HD void ic_bn_cons(const double *eos, const double *V, double *Q)
{
    const double gs = eos[0], gg = eos[1], ps = eos[2], pg = eos[3];
    Q[0] = V[8] * V[0];
    Q[1] = Q[0] * V[1];
    Q[2] = Q[0] * V[2];
    Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) + (V[3] + gs * ps) / (V[0] * (gs - 1.0)));
    Q[4] = (1.0 - V[8]) * V[4];
    Q[5] = Q[4] * V[5];
    Q[6] = Q[4] * V[6];
    Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) + (V[7] + gg * pg) / (V[4] * (gg - 1.0)));
    Q[8] = V[8];
}

// Whether a formula kind belongs to a system (0 Euler, 1 Baer-Nunziato, 2 advection-diffusion).
ADG_HD bool ic_fits(int pde, int kind)
{
    if (pde == 1) return kind == kIcBnCircle || kind == kIcBnDisc || kind == kIcBnRiemann;
    if (pde == 2) return kind == kIcSine;
    return kind == kIcVortex || kind == kIcQuadrant || kind == kIcDmr;
}

// Conserved state at (x, y); gamma (Euler) or eos (Baer-Nunziato).
ADG_HD void ic_field(const IcSpec &ic, double gamma, const double *eos, double x, double y, double *Q)
{
#pragma clang fp contract(off)
    const double *p = ic.p;
    if (ic.kind == kIcSine) {
        const double pi = 3.141592653589793;
        Q[0] = p[0] * sin(2.0 * pi * p[1] * x) * sin(2.0 * pi * p[2] * y);
        return;
    }
    if (ic.kind == kIcBnRiemann) {  // two primitive states split at x = p[0], Baer-Nunziato/initialize.f90:314-403
        ic_bn_cons(eos, p + (x <= p[0] ? 1 : 10), Q);
        return;
    }
    if (ic.kind == kIcBnDisc) {
        const bool inside = (x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) <= p[2] * p[2];
        ic_bn_cons(eos, p + (inside ? 3 : 12), Q);
        return;
    }
    if (ic.kind == kIcBnCircle) {
        const bool inside = (x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) <= p[2] * p[2];
        double V[9];
        V[0] = p[5];
        V[1] = V[2] = 0.0;
        V[3] = inside ? p[3] : p[4];
        V[4] = p[6];
        V[5] = V[6] = 0.0;
        V[7] = V[3];
        V[8] = inside ? p[7] : p[8];
        ic_bn_cons(eos, V, Q);
        return;
    }
    double V[4];
    if (ic.kind == kIcVortex) {
        const double pi = 3.141592653589793;
        const double beta = p[0], dxr = x - p[1], dyr = y - p[2], r2 = dxr * dxr + dyr * dyr;
        const double dT = -(gamma - 1.0) * beta * beta / (8.0 * gamma * pi * pi) * exp(1.0 - r2);
        const double rho = pow(1.0 + dT, 1.0 / (gamma - 1.0));
        const double f = beta / (2.0 * pi) * exp(0.5 * (1.0 - r2));
        V[0] = rho;
        V[1] = p[3] - f * dyr;
        V[2] = p[4] + f * dxr;
        V[3] = pow(rho, gamma);
    } else if (ic.kind == kIcQuadrant) {
        const bool east = x > p[0], north = y > p[1];
        const double *s = p + 2 + 4 * (east ? (north ? 0 : 3) : (north ? 1 : 2));
        for (int v = 0; v < 4; v++) V[v] = s[v];
    } else {  // kIcDmr
        const double *s = x <= p[0] + y / p[1] ? p + 2 : p + 6;
        for (int v = 0; v < 4; v++) V[v] = s[v];
    }
    ic_euler_cons(gamma, V, Q);
}

}  // namespace aderdg
