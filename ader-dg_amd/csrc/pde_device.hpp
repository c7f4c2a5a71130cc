// pde_device.hpp -- pointwise device routines of the two PDE systems and the
// Fortran intrinsic semantics the kernels reproduce.
//
// Association order follows the reference expressions exactly, so that the
// build without floating-point contraction is bit-identical to the reference
// routines (checked by the GPU parity tests).
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

#define ADG_DEV __device__ __forceinline__

namespace aderdg {

// Loads of data that no kernel writes while it runs (solve table, work tables,
// the per-solve state outside k_dt / k_solves, operators) through the constant
// address space: uniform addresses then become scalar loads into SGPRs.
#if defined(__HIP_DEVICE_COMPILE__)
#define ADG_CONST __attribute__((address_space(4)))
#else  // the host pass only type-checks the device code
#define ADG_CONST
#endif
template <class T>
ADG_DEV T ld_const(const T *p)
{
    return *(const ADG_CONST T *)p;
}
// the operator matrices, read-only for the whole run
using OpPtr = const ADG_CONST Operators *;

// Workgroups of a launch go to the chip's 8 XCDs in turn (workgroup k to XCD
// k mod 8), each XCD with its own L2.  Work-table entries next to each other
// touch the same cache lines (neighbouring cells of the per-cell and face
// planes), so each round of gridDim.x entries is dealt out as 8 contiguous
// ranges, one per XCD: workgroup k takes entry (k mod 8) * (G/8) + k / 8 of the
// round (a permutation of the round; the last G mod 8 workgroups keep their own).
constexpr int kXcds = 8;
ADG_DEV int xcd_entry(int blk, int nblk)
{
    const int per = nblk / kXcds;
    return blk < per * kXcds ? (blk % kXcds) * per + blk / kXcds : blk;
}

// Two-argument MAX/MIN as flang evaluates them: (a > b ? a : b), so a NaN
// second argument wins and a NaN first argument loses.
ADG_DEV double f_max(double a, double b) { return a > b ? a : b; }
ADG_DEV double f_min(double a, double b) { return a < b ? a : b; }
// One element of MAXVAL/MINVAL, which skip NaN elements; start from NaN.  IEEE
// maxNum/minNum (v_max_f64 / v_min_f64) skip a NaN operand the same way; the two
// differ only in which zero a fold of +0 and -0 returns, which no comparison or
// difference downstream can see.
ADG_DEV double fold_max(double m, double e) { return __builtin_fmax(m, e); }
ADG_DEV double fold_min(double m, double e) { return __builtin_fmin(m, e); }

// a / b.  The exact build divides (IEEE, the reference's arithmetic); the
// product build multiplies by a Newton-Raphson-refined hardware reciprocal
// (v_rcp_f64, within about 2^-25 relative, and two refinements r <- r + r (1 - b r),
// after which r is within half an ulp of 1 / b): measured on gfx950 over 2^20
// random a, b spanning 40 decades, a r lands within 1.43 ulp of the quotient
// (with one refinement r is within 10.9 ulp and a r within 18.2 ulp).
// Divisions by the same b share r.
ADG_DEV double rdiv(double a, double b)
{
#if defined(ADERDG_EXACT) || !defined(__HIP_DEVICE_COMPILE__)
    return a / b;
#else
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    return a * r;
#endif
}

// a / b where b stays the same over a loop and rb = 1 / b was formed once before it.  The
// exact build divides (the reference's rounding); the product build multiplies by the
// reciprocal, within an ulp or two of the quotient, which takes the division out of the loop.
ADG_DEV double div_inv(double a, double b, double rb)
{
#if defined(ADERDG_EXACT) || !defined(__HIP_DEVICE_COMPILE__)
    (void)rb;
    return a / b;
#else
    (void)b;
    return a * rb;
#endif
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

// Linear advection-diffusion, Advection-Diffusion/pde.f90 (its kernels are in
// ad_kernels.hip; this type only sizes the shared initial-condition kernel).
struct AdPDE {
    static constexpr int V = 1;
    static constexpr bool kNcp = false;
};

// ---------------------------------------------------------------------------
// 2-D Euler, Euler/pde.f90
// ---------------------------------------------------------------------------
struct EulerPDE {
    static constexpr int V = 4;
    static constexpr bool kNcp = false;
    // variables whose flux is identically zero / whose gradient the NCP source reads / whose
    // NCP source is identically -0 (none: no NCP)
    static constexpr bool flux_zero(int) { return false; }
    static constexpr bool grad_used(int) { return false; }
    static constexpr bool src_zero(int) { return false; }
    // x- and y-momentum components (a reflective side negates the normal one)
    static constexpr unsigned kMomX = 1u << 1, kMomY = 1u << 2;

    // PDEFlux, Euler/pde.f90:46-69.  F[0..3] x-flux, F[4..7] y-flux.
    template <class M>
    ADG_DEV static void flux(const M &m, const double *Q, double *F)
    {
        const double g = m.gamma;
        const double irho = rdiv(1.0, Q[0]);
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
    template <class M>
    ADG_DEV static double maxabs_eig(const M &m, const double *Q, double nx, double ny)
    {
        const double g = m.gamma;
        const double u = rdiv(Q[1] * nx + Q[2] * ny, Q[0]);
        const double p = (g - 1.0) * (Q[3] - rdiv(0.5 * (Q[1] * Q[1] + Q[2] * Q[2]), Q[0]));
        const double c = __builtin_sqrt(rdiv(g * p, Q[0]));
        double mx = __builtin_nan("");
        mx = fold_max(mx, __builtin_fabs(u - c));
        mx = fold_max(mx, __builtin_fabs(u));
        mx = fold_max(mx, __builtin_fabs(u));
        mx = fold_max(mx, __builtin_fabs(u + c));
        return mx;
    }

    // RusanovFlux, Euler/pde.f90:97-114.  One flux serves both sides: Dm = Dp.
    template <class M>
    ADG_DEV static void face_flux(const M &m, const double *QL, const double *FL, const double *QR,
                                  const double *FR, double nx, double ny, double *Dm, double *Dp)
    {
        const double smax = f_max(maxabs_eig(m, QL, nx, ny), maxabs_eig(m, QR, nx, ny));
#pragma unroll
        for (int v = 0; v < V; v++) {
            Dm[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * smax * (QR[v] - QL[v]);
            Dp[v] = Dm[v];
        }
    }

    // PDEAssurePositivity, Euler/pde.f90:120-145: true if the state is rejected
    template <class M>
    ADG_DEV static bool inadmissible(const M &m, const double *Q)
    {
        const double irho = rdiv(1.0, Q[0]);
        const double p = (m.gamma - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
        return (Q[0] <= 1.0e-12) || (p <= 1.0e-12);
    }

    ADG_DEV static void ncp(const MeshArgs &, const double *, const double *, double *) {}
};

// ---------------------------------------------------------------------------
// 2-D Baer-Nunziato, Baer-Nunziato/pde.f90 (eos = GAMMA_S, GAMMA_G, PI_S, PI_G)
// Q = (phi_s rho_s, .. u_s, .. v_s, phi_s E_s, phi_g rho_g, .. u_g, .. v_g, phi_g E_g, phi_s)
// ---------------------------------------------------------------------------
struct BnPDE {
    static constexpr int V = 9;
    static constexpr bool kNcp = true;
    // the volume fraction (variable 8) has no flux (F[8] = F[17] = 0), and it is the only
    // variable whose gradient the NCP source reads (ncp below: G[8], G[V + 8])
    static constexpr bool flux_zero(int v) { return v == 8; }
    static constexpr bool grad_used(int v) { return v == 8; }
    // the NCP source of the two densities is -0 (ncp below: src[0] = src[4] = -0.0); adding
    // f * (-0) with f >= 0 leaves any value unchanged, so the predictors skip those terms
    static constexpr bool src_zero(int v) { return v == 0 || v == 4; }
    // x- and y-momentum components of both phases
    static constexpr unsigned kMomX = (1u << 1) | (1u << 5), kMomY = (1u << 2) | (1u << 6);

    // PDEFlux, Baer-Nunziato/pde.f90:74-119.  F[0..8] x-flux, F[9..17] y-flux.
    template <class M>
    ADG_DEV static void flux(const M &m, const double *Q, double *F)
    {
        const double *e = m.eos;
        const double phi_s = Q[8], phi_g = 1.0 - Q[8];
        const double rho_s = rdiv(Q[0], phi_s), rho_g = rdiv(Q[4], phi_g);
        const double u_s = rdiv(Q[1], Q[0]), v_s = rdiv(Q[2], Q[0]);
        const double u_g = rdiv(Q[5], Q[4]), v_g = rdiv(Q[6], Q[4]);
        const double E_s = rdiv(Q[3], phi_s), E_g = rdiv(Q[7], phi_g);
        const double p_s = (e[0] - 1.0) * (E_s - 0.5 * rho_s * (u_s * u_s + v_s * v_s)) - e[0] * e[2];
        const double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
        F[0] = phi_s * rho_s * u_s;
        F[1] = phi_s * (rho_s * u_s * u_s + p_s);
        F[2] = phi_s * rho_s * u_s * v_s;
        F[3] = phi_s * u_s * (E_s + p_s);
        F[4] = phi_g * rho_g * u_g;
        F[5] = phi_g * (rho_g * u_g * u_g + p_g);
        F[6] = phi_g * rho_g * u_g * v_g;
        F[7] = phi_g * u_g * (E_g + p_g);
        F[8] = 0.0;
        F[9] = phi_s * rho_s * v_s;
        F[10] = phi_s * rho_s * u_s * v_s;
        F[11] = phi_s * (rho_s * v_s * v_s + p_s);
        F[12] = phi_s * v_s * (E_s + p_s);
        F[13] = phi_g * rho_g * v_g;
        F[14] = phi_g * rho_g * u_g * v_g;
        F[15] = phi_g * (rho_g * v_g * v_g + p_g);
        F[16] = phi_g * v_g * (E_g + p_g);
        F[17] = 0.0;
    }

    // maxval(abs(Lambda)) of PDEEigenvalues, Baer-Nunziato/pde.f90:125-144 (via PDECons2Prim :8-43)
    template <class M>
    ADG_DEV static double maxabs_eig(const M &m, const double *Q, double nx, double ny)
    {
        const double *e = m.eos;
        const double phi_s = Q[8], phi_g = 1.0 - Q[8];
        const double rho_s = rdiv(Q[0], phi_s), rho_g = rdiv(Q[4], phi_g);
        const double u_s = rdiv(Q[1], Q[0]), v_s = rdiv(Q[2], Q[0]);
        const double u_g = rdiv(Q[5], Q[4]), v_g = rdiv(Q[6], Q[4]);
        const double E_s = rdiv(Q[3], phi_s), E_g = rdiv(Q[7], phi_g);
        const double p_s = (e[0] - 1.0) * (E_s - 0.5 * rho_s * (u_s * u_s + v_s * v_s)) - e[0] * e[2];
        const double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
        const double un_s = u_s * nx + v_s * ny;
        const double un_g = u_g * nx + v_g * ny;
        const double a_s = __builtin_sqrt(rdiv(e[0] * (p_s + e[2]), rho_s));
        const double a_g = __builtin_sqrt(rdiv(e[1] * (p_g + e[3]), rho_g));
        double mx = __builtin_nan("");
        mx = fold_max(mx, __builtin_fabs(un_g - a_g));
        mx = fold_max(mx, __builtin_fabs(un_g + a_g));
        mx = fold_max(mx, __builtin_fabs(un_s - a_s));
        mx = fold_max(mx, __builtin_fabs(un_s + a_s));
        mx = fold_max(mx, __builtin_fabs(un_g));
        mx = fold_max(mx, __builtin_fabs(un_s));
        return mx;
    }

    // p_g, u_s, v_s as in PDEMatrixB / PDENCP, Baer-Nunziato/pde.f90:163-176, 207-219
    template <class M>
    ADG_DEV static double gas_pressure(const M &m, const double *Q, double &u_s, double &v_s)
    {
        const double *e = m.eos;
        const double phi_g = 1.0 - Q[8];
        const double rho_g = rdiv(Q[4], phi_g);
        u_s = rdiv(Q[1], Q[0]);
        v_s = rdiv(Q[2], Q[0]);
        const double u_g = rdiv(Q[5], Q[4]), v_g = rdiv(Q[6], Q[4]);
        const double E_g = rdiv(Q[7], phi_g);
        return (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
    }

    // Column 9 of the Roe matrix: 3-point Gauss-Legendre path integral of PDEMatrixB
    // (RoeMatrix, Baer-Nunziato/pde.f90:296-320; PDEMatrixB :150-186).  The other columns are 0.
    template <class M>
    ADG_DEV static void roe_col9(const M &m, const double *Qa, const double *Qb, double nx, double ny,
                                 double *B9)
    {
        const double s15 = __builtin_sqrt(15.0);
        const double sGP[3] = {0.5 - s15 / 10., 0.5, 0.5 + s15 / 10.};
        const double wGP[3] = {5. / 18.0, 8. / 18., 5. / 18.};
#pragma unroll
        for (int r = 0; r < V; r++) B9[r] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double Q[V];
#pragma unroll
            for (int v = 0; v < V; v++) Q[v] = Qa[v] + sGP[i] * (Qb[v] - Qa[v]);
            double u_s, v_s;
            const double p_g = gas_pressure(m, Q, u_s, v_s);
            const double u_n = nx * u_s + ny * v_s;
            double b[V];
            b[0] = 0.0;
            b[1] = -nx * p_g;
            b[2] = -ny * p_g;
            b[3] = -p_g * u_n;
            b[4] = 0.0;
            b[5] = nx * p_g;
            b[6] = ny * p_g;
            b[7] = p_g * u_n;
            b[8] = u_n;
#pragma unroll
            for (int r = 0; r < V; r++) B9[r] = B9[r] + wGP[i] * b[r];
        }
    }

    // matmul(BRoe, dQ) with a Roe matrix whose only non-zero column is 9: the
    // zero columns add +0 terms, so the running sum starts at +0.
    ADG_DEV static double roe_times(const double *B9, int r, const double *dQ)
    {
        double acc = 0.0;
        acc = acc + B9[r] * dQ[8];
        return acc;
    }

    // RusanovFlux (path-conservative), Baer-Nunziato/pde.f90:326-353
    template <class M>
    ADG_DEV static void face_flux(const M &m, const double *QL, const double *FL, const double *QR,
                                  const double *FR, double nx, double ny, double *Dm, double *Dp)
    {
        const double smax = f_max(maxabs_eig(m, QL, nx, ny), maxabs_eig(m, QR, nx, ny));
        double B9[V], dQ[V];
        roe_col9(m, QL, QR, nx, ny, B9);
#pragma unroll
        for (int v = 0; v < V; v++) dQ[v] = QR[v] - QL[v];
#pragma unroll
        for (int v = 0; v < V; v++) {
            const double d = 0.5 * (FL[v] + FR[v]) - 0.5 * smax * dQ[v];
            const double t = roe_times(B9, v, dQ);
            Dm[v] = d - 0.5 * t;
            Dp[v] = d + 0.5 * t;
        }
    }

    // PDEAssurePositivity, Baer-Nunziato/pde.f90:237-266: true if the state is rejected
    template <class M>
    ADG_DEV static bool inadmissible(const M &m, const double *Q)
    {
        const double *e = m.eos;
        const double phi_s = Q[8], phi_g = 1.0 - Q[8];
        const double rho_s = rdiv(Q[0], phi_s), rho_g = rdiv(Q[4], phi_g);
        const double u_s = rdiv(Q[1], Q[0]), v_s = rdiv(Q[2], Q[0]);
        const double u_g = rdiv(Q[5], Q[4]), v_g = rdiv(Q[6], Q[4]);
        const double E_s = rdiv(Q[3], phi_s), E_g = rdiv(Q[7], phi_g);
        const double p_s = (e[0] - 1.0) * (E_s - 0.5 * rho_s * (u_s * u_s + v_s * v_s)) - e[0] * e[2];
        const double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
        return rho_s <= 1.0e-12 || rho_g <= 1.0e-12 || (p_s + e[2]) <= 1.0e-12 || (p_g + e[3]) <= 1.0e-12 ||
               phi_s <= 0.0 || phi_s >= 1.0;
    }

    // PDEFusedSrcNCP = -PDENCP, Baer-Nunziato/pde.f90:277-290, 192-231.  Gx, Gy: gradient components.
    template <class M>
    ADG_DEV static void ncp(const M &m, const double *Q, const double *G, double *src)
    {
        double u_s, v_s;
        const double p_g = gas_pressure(m, Q, u_s, v_s);
        const double qx = G[8], qy = G[V + 8];
        const double un = u_s * qx + v_s * qy;
        src[0] = -0.0;
        src[1] = -(-p_g * qx);
        src[2] = -(-p_g * qy);
        src[3] = -(-p_g * un);
        src[4] = -0.0;
        src[5] = -(p_g * qx);
        src[6] = -(p_g * qy);
        src[7] = -(p_g * un);
        src[8] = -un;
    }
};

// Ghost face state and normal flux of a non-periodic boundary side (codes of
// MeshArgs::side): inflow takes m.qin and its flux along (nx, ny) (the DMR inflow
// arithmetic, Euler/riemann_problem.f90:186-193); transmissive the inside state
// (Baer-Nunziato/riemann_problem.f90:26-34); reflective the inside state with the
// normal momentum of every phase negated and every other flux component negated
// (the DMR wall, Euler/riemann_problem.f90:225-228).
template <class P, class M>
ADG_DEV void side_ghost(const M &m, int code, double nx, double ny, const double *Qi, const double *Fi, double *Q,
                        double *F)
{
    constexpr int V = P::V;
    if (code == kSideInflow) {
        double Fbc[2 * V];
#pragma unroll
        for (int v = 0; v < V; v++) Q[v] = m.qin[v];
        P::flux(m, Q, Fbc);
#pragma unroll
        for (int v = 0; v < V; v++) F[v] = 0.0 + Fbc[v] * nx + Fbc[V + v] * ny;
        return;
    }
    const unsigned mom = code == kSideReflect ? (nx != 0.0 ? P::kMomX : P::kMomY) : 0u;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const bool normal = (mom >> v) & 1u;
        Q[v] = normal ? -Qi[v] : Qi[v];
        F[v] = (mom && !normal) ? -Fi[v] : Fi[v];
    }
}

// A ghost subcell of the subcell recompute outside side gx (-1 left, +1 right, 0
// none) and gy (-1 bottom, +1 top), taken from the clamped or wrapped neighbour
// table: inflow replaces it by m.qin, reflective negates its normal momentum; x
// sides first (the DMR ghosts, Euler/limiter.f90:491-522).  Variable v of the
// subcell is val[v * stride].
template <class P, class M>
ADG_DEV void side_subcell(const M &m, int gx, int gy, double *val, int stride = 1)
{
    constexpr int V = P::V;
    for (int a = 0; a < 2; a++) {
        const int g = a == 0 ? gx : gy;
        if (g == 0) continue;
        const int code = m.side[a == 0 ? (g < 0 ? 0 : 1) : (g < 0 ? 2 : 3)];
        const unsigned mom = a == 0 ? P::kMomX : P::kMomY;
        if (code == kSideInflow) {
#pragma unroll
            for (int v = 0; v < V; v++) val[v * stride] = m.qin[v];
        } else if (code == kSideReflect) {
#pragma unroll
            for (int v = 0; v < V; v++)
                if ((mom >> v) & 1u) val[v * stride] = -val[v * stride];
        }
    }
}

}  // namespace aderdg
