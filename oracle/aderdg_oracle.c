/*
 * aderdg_oracle.c -- CPU restatement of the reference ADER-DG element step
 * (2-D Euler).  TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never
 * linked into the product.
 *
 * Build with -O2 -ffp-contract=off: every expression keeps the association and
 * summation order of the Fortran it restates (matmul sums run over the inner
 * index in increasing order from 0.0, Fortran's left-to-right evaluation of
 * equal-precedence operators), so results are bit-identical to the reference
 * built with flang -O3 -fdefault-real-8 on x86-64 (no FMA).  The CPU tests
 * check that bit-identity routine by routine against oracle/_ref.
 */
#include "aderdg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double g_ops;
#define OPS(k) (g_ops += (double)(k))
void orc_reset_opcount(void) { g_ops = 0.0; }
double orc_get_opcount(void) { return g_ops; }

#define NMAX 10
#define SMAX (2 * NMAX + 1)
#define VMAX 9

struct orc_ctx {
    int pde, N, n, V, s, imax, jmax, bc_mode;
    double xl[2], xr[2], dx[2];
    double gamma;
    double eos[4];  /* Baer-Nunziato: GAMMA_S, GAMMA_G, PI_S, PI_G (Baer-Nunziato/initialize.f90:23-26) */
    double cflpnpm; /* CFL*PNPMTable(N), Euler/calc_time_step.f90:41 */
    double xi[NMAX], w[NMAX], mm[NMAX * NMAX], imm[NMAX * NMAX], kxi[NMAX * NMAX], dudx[NMAX * NMAX];
    double k1[NMAX * NMAX], ik1[NMAX * NMAX], f0[NMAX], flc[NMAX], frc[NMAX];
    double u2l[SMAX * NMAX], l2u[NMAX * SMAX], u2b[NMAX * NMAX], xlob[NMAX], wlob[NMAX];
    double dt, time, tend;
    /* state */
    double *uh, *olduh, *qhi, *Fhi, *qbnd, *Fbnd, *duh, *Fu, *Gu;
    double *Shi, *Fp, *Gp; /* Baer-Nunziato: time-averaged NCP source and the D+ fluctuations (Fu/Gu hold D-) */
    int *recompute, *status, *oldstatus;
    double *Lh, *NewLh, *lmin, *lmax, *lmin_old, *lmax_old;
    /* advection-diffusion: EQN%cx, EQN%cy, EQN%a, bC(1..4), predictor flux variant */
    double ad_c[2], ad_a;
    int ad_bc[4], ad_fixed;
    double *gradQbnd; /* (nVar, nDim, 4, n) per cell */
    double *margin;   /* when recording: the DMP margin of each cell's candidate in the last step */
    int side[4];      /* boundary code per side (orc_set_sides); from bc_mode otherwise */
    double qin[VMAX]; /* conserved inflow state of code-0 sides */
};

/* PNPMTable of Euler/ader_dg.f90:28 */
static const double PNPM[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};

/* ------------------------------------------------------------------ */
/* quadrature, basis, linear algebra (setup)                            */
/* ------------------------------------------------------------------ */

/* gauleg: Euler/quadrature.f90:8-44 */
static void gauleg(double x1, double x2, double *x, double *w, int n)
{
    const double EPS = 1.0e-15, PI = 4.0 * atan(1.0);
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(PI * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > EPS);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* gaulob: Euler/quadrature.f90:47-101 */
static void gaulob(double x1, double x2, double *x, double *w, int n1)
{
    const double EPS = 1.0e-15, PI = 4.0 * atan(1.0);
    int N = n1 - 1;
    double xold[NMAX], P[NMAX][NMAX + 1]; /* P[point][k], k = 1..n1 */
    for (int i = 0; i <= N; i++) x[i] = cos(PI * (double)i / (double)N);
    for (int iter = 1; iter <= 100000; iter++) {
        double dmax = 0.0;
        for (int a = 0; a < n1; a++) xold[a] = x[a];
        for (int a = 0; a < n1; a++) {
            P[a][1] = 1.0;
            P[a][2] = x[a];
            for (int k = 2; k <= N; k++)
                P[a][k + 1] = ((double)(2 * k - 1) * x[a] * P[a][k] - (double)(k - 1) * P[a][k - 1]) / (double)k;
        }
        for (int a = 0; a < n1; a++) x[a] = xold[a] - (x[a] * P[a][n1] - P[a][N]) / ((double)n1 * P[a][n1]);
        for (int a = 0; a < n1; a++) {
            double d = fabs(x[a] - xold[a]);
            if (d > dmax) dmax = d;
        }
        if (dmax <= EPS) break;
    }
    for (int a = 0; a < n1; a++) {
        w[a] = 2.0 / ((double)(N * n1) * (P[a][n1] * P[a][n1]));
        w[a] = 0.5 * w[a] * (x2 - x1);
        x[a] = x1 + 0.5 * (x[a] + 1.0) * (x2 - x1);
    }
}

/* basis_func_1d: Euler/basis.f90:9-41 (Lagrange basis through xin) */
static void basis_1d(const double *xin, int n, double xi, double *phi, double *phi_xi)
{
    for (int m = 0; m < n; m++) {
        phi[m] = 1.0;
        phi_xi[m] = 0.0;
    }
    for (int m = 0; m < n; m++) {
        for (int j = 0; j < n; j++) {
            if (j == m) continue;
            phi[m] = phi[m] * (xi - xin[j]) / (xin[m] - xin[j]);
        }
        for (int i = 0; i < n; i++) {
            if (i == m) continue;
            double tmp = 1.0;
            for (int j = 0; j < n; j++) {
                if (j == i || j == m) continue;
                tmp = tmp * (xi - xin[j]) / (xin[m] - xin[j]);
            }
            phi_xi[m] = phi_xi[m] + tmp / (xin[m] - xin[i]);
        }
    }
}

/* MatrixInverse: Euler/la.f90:8-65 (Gauss-Jordan on the transposed matrix,
 * column pivoting by maxloc).  A and iA column-major NxN. */
static int matinv(int N, const double *A, double *iA)
{
    double C[2 * (NMAX + 2)][NMAX + 2]; /* C[row][col], rows 0..2N-1 */
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) C[r][c] = A[c + N * r]; /* C(1:N,:) = transpose(A) */
    for (int r = N; r < 2 * N; r++)
        for (int c = 0; c < N; c++) C[r][c] = 0.0;
    for (int i = 0; i < N; i++) C[N + i][i] = 1.0;
    for (int i = 0; i < N; i++) {
        int j = i;
        double best = fabs(C[i][i]);
        for (int c = i + 1; c < N; c++)
            if (fabs(C[i][c]) > best) {
                best = fabs(C[i][c]);
                j = c;
            }
        for (int r = 0; r < 2 * N; r++) {
            double t = C[r][j];
            C[r][j] = C[r][i];
            C[r][i] = t;
        }
        if (fabs(C[i][i]) <= 1.0e-15) return -1;
        double piv = 1.0 / C[i][i];
        for (int r = 0; r < 2 * N; r++) C[r][i] = C[r][i] * piv;
        for (int c = i + 1; c < N; c++) {
            double f = C[i][c];
            for (int r = 0; r < 2 * N; r++) C[r][c] = C[r][c] - f * C[r][i];
        }
    }
    for (int i = N - 1; i >= 0; i--)
        for (int c = i - 1; c >= 0; c--) {
            double f = C[i][c];
            for (int r = 0; r < 2 * N; r++) C[r][c] = C[r][c] - f * C[r][i];
        }
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) iA[r + N * c] = C[N + c][r]; /* iA = transpose(C(N+1:2N,:)) */
    return 0;
}

/* operator construction: Euler/initialize.f90:80-213 */
static void build_ops(orc_ctx *c)
{
    int n = c->n, s = c->s;
    double phi[NMAX], phx[NMAX], phi0[NMAX], phi1[NMAX];
    gauleg(0.0, 1.0, c->xi, c->w, n);
    memset(c->mm, 0, sizeof c->mm);
    memset(c->kxi, 0, sizeof c->kxi);
    for (int g = 0; g < n; g++) {
        basis_1d(c->xi, n, c->xi[g], phi, phx);
        for (int k = 0; k < n; k++)
            for (int l = 0; l < n; l++) {
                c->mm[k + n * l] = c->mm[k + n * l] + c->w[g] * phi[k] * phi[l];
                c->kxi[k + n * l] = c->kxi[k + n * l] + c->w[g] * phx[k] * phi[l];
            }
    }
    if (matinv(n, c->mm, c->imm)) fprintf(stderr, "orc: singular mass matrix\n");
    for (int a = 0; a < n; a++) /* dudx = matmul(iMM, transpose(Kxi)) */
        for (int b = 0; b < n; b++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc = acc + c->imm[a + n * k] * c->kxi[b + n * k];
            c->dudx[a + n * b] = acc;
        }
    basis_1d(c->xi, n, 0.0, phi0, phx);
    basis_1d(c->xi, n, 1.0, phi1, phx);
    for (int k = 0; k < n; k++) {
        c->f0[k] = phi0[k];
        c->flc[k] = phi0[k];
        c->frc[k] = phi1[k];
        for (int l = 0; l < n; l++) c->k1[k + n * l] = phi1[k] * phi1[l] - c->kxi[k + n * l];
    }
    if (matinv(n, c->k1, c->ik1)) fprintf(stderr, "orc: singular K1\n");
    /* subcell maps, Euler/initialize.f90:175-201 */
    double dxi = 1.0 / (double)s;
    memset(c->u2l, 0, sizeof c->u2l);
    for (int i = 1; i <= s; i++) {
        double xi1 = (double)(i - 1) * dxi;
        for (int ii = 0; ii < n; ii++) {
            double x = xi1 + dxi * c->xi[ii];
            basis_1d(c->xi, n, x, phi, phx);
            for (int a = 0; a < n; a++) c->u2l[(i - 1) + s * a] = c->u2l[(i - 1) + s * a] + c->w[ii] * phi[a];
        }
    }
    int m = n + 1;
    double lsq[(NMAX + 1) * (NMAX + 1)], ilsq[(NMAX + 1) * (NMAX + 1)], rhs[(NMAX + 1) * SMAX];
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            double acc = 0.0;
            for (int p = 0; p < s; p++) acc = acc + c->u2l[p + s * a] * c->u2l[p + s * b];
            lsq[a + m * b] = 2.0 * acc;
        }
    for (int a = 0; a < n; a++) {
        lsq[n + m * a] = c->w[a];
        lsq[a + m * n] = -c->w[a];
    }
    lsq[n + m * n] = 0.0;
    if (matinv(m, lsq, ilsq)) fprintf(stderr, "orc: singular LSQ matrix\n");
    for (int a = 0; a < n; a++)
        for (int p = 0; p < s; p++) rhs[a + m * p] = 2.0 * c->u2l[p + s * a];
    for (int p = 0; p < s; p++) rhs[n + m * p] = dxi;
    for (int a = 0; a < n; a++)
        for (int p = 0; p < s; p++) {
            double acc = 0.0;
            for (int k = 0; k < m; k++) acc = acc + ilsq[a + m * k] * rhs[k + m * p];
            c->l2u[a + n * p] = acc;
        }
    gaulob(0.0, 1.0, c->xlob, c->wlob, n);
    for (int ii = 0; ii < n; ii++) {
        basis_1d(c->xi, n, c->xlob[ii], phi, phx);
        for (int a = 0; a < n; a++) c->u2b[ii + n * a] = phi[a];
    }
}

/* ------------------------------------------------------------------ */
/* PDE pointwise routines: Euler/pde.f90                                */
/* ------------------------------------------------------------------ */

/* PDEFlux: Euler/pde.f90:46-69.  F(V,2) column-major. */
static void euler_flux(double g, const double *Q, double *F)
{
    double irho = 1.0 / Q[0];
    double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
    F[0] = Q[1];
    F[1] = irho * Q[1] * Q[1] + p;
    F[2] = irho * Q[1] * Q[2];
    F[3] = irho * Q[1] * (Q[3] + p);
    F[4] = Q[2];
    F[5] = irho * Q[2] * Q[1];
    F[6] = irho * Q[2] * Q[2] + p;
    F[7] = irho * Q[2] * (Q[3] + p);
    OPS(1 + 9 + 3 + 2 + 3 + 2 + 3 + 3);
}

/* PDEEigenvalues: Euler/pde.f90:75-91 */
static void euler_eigen(double g, const double *Q, const double *nv, double *L)
{
    double u = (Q[1] * nv[0] + Q[2] * nv[1]) / Q[0];
    double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]);
    double cs = sqrt(g * p / Q[0]);
    L[0] = u - cs;
    L[1] = u;
    L[2] = u;
    L[3] = u + cs;
    OPS(4 + 8 + 3 + 2);
}

/* Fortran intrinsic semantics as compiled by flang (checked on a probe program):
 * MAXVAL/MINVAL skip NaN elements (NaN only when every element is NaN), while the
 * two-argument MAX(a,b)/MIN(a,b) evaluate as (a > b ? a : b) / (a < b ? a : b),
 * so a NaN second argument wins.  Multi-argument MIN/MAX pair up left to right. */
static double f_max(double a, double b) { return a > b ? a : b; }
static double f_min(double a, double b) { return a < b ? a : b; }
static double maxval_s(const double *x, int cnt, int stride)
{
    double m = NAN;
    for (int k = 0; k < cnt; k++) {
        double e = x[(size_t)k * stride];
        if (isnan(e)) continue;
        if (isnan(m) || e > m) m = e;
    }
    return m;
}
static double minval_s(const double *x, int cnt, int stride)
{
    double m = NAN;
    for (int k = 0; k < cnt; k++) {
        double e = x[(size_t)k * stride];
        if (isnan(e)) continue;
        if (isnan(m) || e < m) m = e;
    }
    return m;
}

/* maxval(abs(L)) */
static double maxabs(const double *L, int V)
{
    double a[VMAX];
    for (int v = 0; v < V; v++) a[v] = fabs(L[v]);
    return maxval_s(a, V, 1);
}

/* RusanovFlux: Euler/pde.f90:97-114 */
static void euler_rusanov(double g, const double *QL, const double *FL, const double *QR, const double *FR,
                          const double *nv, double *flux)
{
    double LL[4], LR[4];
    euler_eigen(g, QL, nv, LL);
    euler_eigen(g, QR, nv, LR);
    double smax = f_max(maxabs(LL, 4), maxabs(LR, 4));
    for (int v = 0; v < 4; v++) flux[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * smax * (QR[v] - QL[v]);
    OPS(4 * 6);
}

/* PDEAssurePositivity: Euler/pde.f90:120-145 (returns iErr) */
static int euler_positivity(double g, const double *Q)
{
    int ierr = 0;
    double irho = 1.0 / Q[0];
    double p = (g - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
    if (Q[0] <= 1.0e-12) ierr = -1;
    if (p <= 1.0e-12) ierr = -2;
    OPS(1 + 8);
    return ierr;
}

/* PDEPrim2Cons: Euler/pde.f90:29-40 */
static void euler_prim2cons(double g, const double *Vp, double *Q)
{
    Q[0] = Vp[0];
    Q[1] = Vp[0] * Vp[1];
    Q[2] = Vp[0] * Vp[2];
    Q[3] = Vp[3] / (g - 1.0) + 0.5 * Vp[0] * (Vp[1] * Vp[1] + Vp[2] * Vp[2]);
}


/* ------------------------------------------------------------------ */
/* Baer-Nunziato pointwise routines: Baer-Nunziato/pde.f90             */
/* Q = (phi_s rho_s, .. u_s, .. v_s, phi_s E_s, phi_g rho_g, .. u_g, .. v_g, phi_g E_g, phi_s) */
/* ------------------------------------------------------------------ */

/* PDECons2Prim: Baer-Nunziato/pde.f90:8-43 */
static void bn_cons2prim(const double *e, const double *Q, double *V)
{
    double phi_s = Q[8], phi_g = 1.0 - Q[8];
    double rho_s = Q[0] / phi_s, rho_g = Q[4] / phi_g;
    double u_s = Q[1] / Q[0], v_s = Q[2] / Q[0];
    double u_g = Q[5] / Q[4], v_g = Q[6] / Q[4];
    double E_s = Q[3] / phi_s, E_g = Q[7] / phi_g;
    double p_s = (e[0] - 1.0) * (E_s - 0.5 * rho_s * (u_s * u_s + v_s * v_s)) - e[0] * e[2];
    double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
    V[0] = rho_s; V[1] = u_s; V[2] = v_s; V[3] = p_s;
    V[4] = rho_g; V[5] = u_g; V[6] = v_g; V[7] = p_g; V[8] = phi_s;
    OPS(8 + 2 * 11);
}

/* PDEPrim2Cons: Baer-Nunziato/pde.f90:49-68 */
static void bn_prim2cons(const double *e, const double *V, double *Q)
{
    Q[0] = V[8] * V[0];
    Q[1] = Q[0] * V[1];
    Q[2] = Q[0] * V[2];
    Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) + (V[3] + e[0] * e[2]) / (V[0] * (e[0] - 1.0)));
    Q[4] = (1.0 - V[8]) * V[4];
    Q[5] = Q[4] * V[5];
    Q[6] = Q[4] * V[6];
    Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) + (V[7] + e[1] * e[3]) / (V[4] * (e[1] - 1.0)));
    Q[8] = V[8];
}

/* PDEFlux: Baer-Nunziato/pde.f90:74-119.  F(9,2) column-major. */
static void bn_flux(const double *e, const double *Q, double *F)
{
    double phi_s = Q[8], phi_g = 1.0 - Q[8];
    double rho_s = Q[0] / phi_s, rho_g = Q[4] / phi_g;
    double u_s = Q[1] / Q[0], v_s = Q[2] / Q[0];
    double u_g = Q[5] / Q[4], v_g = Q[6] / Q[4];
    double E_s = Q[3] / phi_s, E_g = Q[7] / phi_g;
    double p_s = (e[0] - 1.0) * (E_s - 0.5 * rho_s * (u_s * u_s + v_s * v_s)) - e[0] * e[2];
    double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
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
    OPS(2 + 8 + 2 * 9 + 2 * 3 * 8);
}

/* PDEEigenvalues: Baer-Nunziato/pde.f90:125-144 */
static void bn_eigen(const double *e, const double *Q, const double *nv, double *L)
{
    double V[9];
    bn_cons2prim(e, Q, V);
    double u_s = V[1] * nv[0] + V[2] * nv[1];
    double u_g = V[5] * nv[0] + V[6] * nv[1];
    double a_s = sqrt(e[0] * (V[3] + e[2]) / V[0]);
    double a_g = sqrt(e[1] * (V[7] + e[3]) / V[4]);
    L[0] = u_g - a_g; L[1] = u_g + a_g; L[2] = u_s - a_s; L[3] = u_s + a_s;
    L[4] = u_g; L[5] = u_g; L[6] = u_s; L[7] = u_s; L[8] = u_s;
    OPS(6 + 8 + 4);
}

/* p_g and the normal solid velocity shared by PDEMatrixB and PDENCP */
static double bn_gas_pressure(const double *e, const double *Q, double *u_s, double *v_s)
{
    double phi_g = 1.0 - Q[8];
    double rho_g = Q[4] / phi_g;
    *u_s = Q[1] / Q[0];
    *v_s = Q[2] / Q[0];
    double u_g = Q[5] / Q[4], v_g = Q[6] / Q[4];
    double E_g = Q[7] / phi_g;
    OPS(1 + 6 + 11);
    return (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
}

/* PDEMatrixB: Baer-Nunziato/pde.f90:150-186.  Bn(9,9) column-major; only column 9 is non-zero. */
static void bn_matrix_b(const double *e, const double *Q, const double *nv, double *Bn)
{
    double u_s, v_s;
    for (int k = 0; k < 81; k++) Bn[k] = 0.0;
    double phi_g = 1.0 - Q[8];
    double rho_g = Q[4] / phi_g;
    u_s = Q[1] / Q[0];
    v_s = Q[2] / Q[0];
    double u_n = nv[0] * u_s + nv[1] * v_s;
    double u_g = Q[5] / Q[4], v_g = Q[6] / Q[4];
    double E_g = Q[7] / phi_g;
    double p_g = (e[1] - 1.0) * (E_g - 0.5 * rho_g * (u_g * u_g + v_g * v_g)) - e[1] * e[3];
    Bn[1 + 9 * 8] = -nv[0] * p_g;
    Bn[2 + 9 * 8] = -nv[1] * p_g;
    Bn[3 + 9 * 8] = -p_g * u_n;
    Bn[5 + 9 * 8] = nv[0] * p_g;
    Bn[6 + 9 * 8] = nv[1] * p_g;
    Bn[7 + 9 * 8] = p_g * u_n;
    Bn[8 + 9 * 8] = u_n;
    OPS(1 + 6 + 3 + 11 + 6);
}

/* PDEFusedSrcNCP = -PDENCP: Baer-Nunziato/pde.f90:277-290, 192-231.  gradQ(9,2) column-major. */
static void bn_fused_src_ncp(const double *e, const double *Q, const double *gradQ, double *src)
{
    double u_s, v_s;
    double p_g = bn_gas_pressure(e, Q, &u_s, &v_s);
    const double *Qx = gradQ, *Qy = gradQ + 9;
    double ncp[9];
    ncp[0] = 0.0;
    ncp[1] = -p_g * Qx[8];
    ncp[2] = -p_g * Qy[8];
    ncp[3] = -p_g * (u_s * Qx[8] + v_s * Qy[8]);
    ncp[4] = 0.0;
    ncp[5] = p_g * Qx[8];
    ncp[6] = p_g * Qy[8];
    ncp[7] = p_g * (u_s * Qx[8] + v_s * Qy[8]);
    ncp[8] = u_s * Qx[8] + v_s * Qy[8];
    for (int v = 0; v < 9; v++) src[v] = -ncp[v];
    OPS(3 * 4 + 4);
}

/* RoeMatrix: Baer-Nunziato/pde.f90:296-320 (3-point Gauss-Legendre path integral) */
static void bn_roe_matrix(const double *e, const double *Qa, const double *Qb, const double *nv, double *BRoe)
{
    const double sGP[3] = {0.5 - sqrt(15.0) / 10., 0.5, 0.5 + sqrt(15.) / 10.};
    const double wGP[3] = {5. / 18.0, 8. / 18., 5. / 18.};
    double B[81], Q[9];
    for (int k = 0; k < 81; k++) BRoe[k] = 0.0;
    for (int i = 0; i < 3; i++) {
        for (int v = 0; v < 9; v++) Q[v] = Qa[v] + sGP[i] * (Qb[v] - Qa[v]);
        bn_matrix_b(e, Q, nv, B);
        for (int k = 0; k < 81; k++) BRoe[k] = BRoe[k] + wGP[i] * B[k];
        OPS(9 * 3 + 7 * 2);
    }
}

/* RusanovFlux (path-conservative): Baer-Nunziato/pde.f90:326-353 */
static void bn_rusanov(const double *e, const double *QL, const double *FL, const double *QR, const double *FR,
                       const double *nv, double *Dm, double *Dp)
{
    double LL[9], LR[9], Bn[81], temp[9];
    bn_eigen(e, QL, nv, LL);
    bn_eigen(e, QR, nv, LR);
    double smax = f_max(maxabs(LL, 9), maxabs(LR, 9));
    bn_roe_matrix(e, QL, QR, nv, Bn);
    for (int v = 0; v < 9; v++) Dp[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * smax * (QR[v] - QL[v]);
    for (int r = 0; r < 9; r++) {
        double acc = 0.0;
        for (int k = 0; k < 9; k++) acc = acc + Bn[r + 9 * k] * (QR[k] - QL[k]);
        temp[r] = acc;
    }
    for (int v = 0; v < 9; v++) {
        Dm[v] = Dp[v] - 0.5 * temp[v];
        Dp[v] = Dp[v] + 0.5 * temp[v];
    }
    OPS(9 * 6 + 9 * 3 + 9 * 4);
}

/* PDEAssurePositivity: Baer-Nunziato/pde.f90:237-266 (returns iErr) */
static int bn_positivity(const double *e, const double *Q)
{
    double V[9];
    int ierr = 0;
    bn_cons2prim(e, Q, V);
    if (V[0] <= 1.0e-12 || V[4] <= 1.0e-12) ierr = -1;
    if ((V[3] + e[2]) <= 1.0e-12 || (V[7] + e[3]) <= 1.0e-12) ierr = -2;
    if (V[8] <= 0.0 || V[8] >= 1.0) ierr = -3;
    OPS(2);
    return ierr;
}

static void pde_flux(const orc_ctx *c, const double *Q, double *F)
{
    if (c->pde == ORC_PDE_BN)
        bn_flux(c->eos, Q, F);
    else
        euler_flux(c->gamma, Q, F);
}

static void pde_eigen(const orc_ctx *c, const double *Q, const double *nv, double *L)
{
    if (c->pde == ORC_PDE_BN)
        bn_eigen(c->eos, Q, nv, L);
    else
        euler_eigen(c->gamma, Q, nv, L);
}

static int pde_positivity(const orc_ctx *c, const double *Q)
{
    return c->pde == ORC_PDE_BN ? bn_positivity(c->eos, Q) : euler_positivity(c->gamma, Q);
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */

orc_ctx *orc_create(int pde, int N, int imax, int jmax, const double xl[2], const double xr[2], int bc_mode,
                    const double *eos)
{
    if ((pde != ORC_PDE_EULER && pde != ORC_PDE_BN && pde != ORC_PDE_AD) || N < 1 || N > 9 || imax < 1 || jmax < 1)
        return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->pde = pde;
    c->N = N;
    c->n = N + 1;
    c->s = 2 * N + 1;
    c->V = pde == ORC_PDE_BN ? 9 : (pde == ORC_PDE_AD ? 1 : 4);
    c->imax = imax;
    c->jmax = jmax;
    c->bc_mode = bc_mode;
    for (int k = 0; k < 4; k++) c->side[k] = bc_mode == ORC_BC_PERIODIC ? 3 : 1;
    c->gamma = eos[0];
    if (pde == ORC_PDE_BN)
        for (int k = 0; k < 4; k++) c->eos[k] = eos[k];
    for (int d = 0; d < 2; d++) {
        c->xl[d] = xl[d];
        c->xr[d] = xr[d];
    }
    c->dx[0] = (xr[0] - xl[0]) / (double)imax; /* dx = (xR-xL)/VMAX, Euler/initialize.f90:43 */
    c->dx[1] = (xr[1] - xl[1]) / (double)jmax;
    c->cflpnpm = 0.9 * PNPM[N];
    build_ops(c);
    size_t nc = (size_t)imax * jmax, V = c->V, n = c->n, s = c->s;
    c->uh = (double *)calloc(V * n * n * nc, sizeof(double));
    c->olduh = (double *)calloc(V * n * n * nc, sizeof(double));
    c->qhi = (double *)calloc(V * n * n * nc, sizeof(double));
    c->Fhi = (double *)calloc(2 * V * n * n * nc, sizeof(double));
    c->duh = (double *)calloc(V * n * n * nc, sizeof(double));
    c->qbnd = (double *)calloc(4 * V * n * nc, sizeof(double));
    c->Fbnd = (double *)calloc(4 * V * n * nc, sizeof(double));
    c->Fu = (double *)calloc(V * n * (imax + 1) * jmax, sizeof(double));
    c->Gu = (double *)calloc(V * n * imax * (jmax + 1), sizeof(double));
    c->Fp = (double *)calloc(V * n * (imax + 1) * jmax, sizeof(double));
    c->Gp = (double *)calloc(V * n * imax * (jmax + 1), sizeof(double));
    c->Shi = (double *)calloc(V * n * n * nc, sizeof(double));
    c->recompute = (int *)calloc(nc, sizeof(int));
    c->status = (int *)calloc(nc, sizeof(int));
    c->oldstatus = (int *)calloc(nc, sizeof(int));
    c->Lh = (double *)calloc(V * s * s * nc, sizeof(double));
    c->NewLh = (double *)calloc(V * s * s * nc, sizeof(double));
    c->lmin = (double *)calloc(V * nc, sizeof(double));
    c->lmax = (double *)calloc(V * nc, sizeof(double));
    c->lmin_old = (double *)calloc(V * nc, sizeof(double));
    c->lmax_old = (double *)calloc(V * nc, sizeof(double));
    c->gradQbnd = (double *)calloc(2 * 4 * V * n * nc, sizeof(double));
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->uh); free(c->olduh); free(c->qhi); free(c->Fhi); free(c->duh); free(c->qbnd); free(c->Fbnd);
    free(c->Fu); free(c->Gu); free(c->recompute); free(c->status); free(c->oldstatus); free(c->Lh);
    free(c->NewLh); free(c->lmin); free(c->lmax); free(c->lmin_old); free(c->lmax_old);
    free(c->Fp); free(c->Gp); free(c->Shi); free(c->gradQbnd); free(c->margin);
    free(c);
}

void orc_get_ops(const orc_ctx *c, double *xi, double *w, double *mm, double *kxi, double *dudx, double *ik1,
                 double *f0, double *flc, double *frc, double *u2l, double *l2u, double *u2b, double *xlob,
                 double *wlob, double *k1)
{
    int n = c->n, s = c->s;
    memcpy(xi, c->xi, n * sizeof(double));
    memcpy(w, c->w, n * sizeof(double));
    memcpy(mm, c->mm, n * n * sizeof(double));
    memcpy(kxi, c->kxi, n * n * sizeof(double));
    memcpy(dudx, c->dudx, n * n * sizeof(double));
    memcpy(ik1, c->ik1, n * n * sizeof(double));
    memcpy(f0, c->f0, n * sizeof(double));
    memcpy(flc, c->flc, n * sizeof(double));
    memcpy(frc, c->frc, n * sizeof(double));
    memcpy(u2l, c->u2l, s * n * sizeof(double));
    memcpy(l2u, c->l2u, n * s * sizeof(double));
    memcpy(u2b, c->u2b, n * n * sizeof(double));
    memcpy(xlob, c->xlob, n * sizeof(double));
    memcpy(wlob, c->wlob, n * sizeof(double));
    memcpy(k1, c->k1, n * n * sizeof(double));
}

void orc_set_scalars(orc_ctx *c, double dt, double time, double tend)
{
    c->dt = dt;
    c->time = time;
    c->tend = tend;
}

void orc_get_scalars(const orc_ctx *c, double *dt, double *time, double *tend)
{
    *dt = c->dt;
    *time = c->time;
    *tend = c->tend;
}

static size_t ncell(const orc_ctx *c) { return (size_t)c->imax * c->jmax; }

/* Test support: record the DMP margin (orc_dmp_margin, < 0 where the cell is troubled) of
 * every cell's candidate solution during each step; orc_get_margins copies the last step's. */
void orc_record_margins(orc_ctx *c, int on)
{
    free(c->margin);
    c->margin = on ? (double *)calloc(ncell(c), sizeof(double)) : NULL;
}

void orc_get_margins(const orc_ctx *c, double *out)
{
    if (c->margin) memcpy(out, c->margin, ncell(c) * sizeof(double));
}


void orc_set_uh(orc_ctx *c, const double *uh) { memcpy(c->uh, uh, c->V * c->n * c->n * ncell(c) * sizeof(double)); }
void orc_get_uh(const orc_ctx *c, double *uh) { memcpy(uh, c->uh, c->V * c->n * c->n * ncell(c) * sizeof(double)); }

void orc_set_limiter(orc_ctx *c, const int *status, const double *lh)
{
    size_t blk = (size_t)c->V * c->s * c->s;
    for (size_t k = 0; k < ncell(c); k++) {
        c->status[k] = c->oldstatus[k] = status[k] ? 1 : 0;
        c->recompute[k] = 0;
        if (status[k]) {
            memcpy(c->Lh + k * blk, lh + k * blk, blk * sizeof(double));
            memcpy(c->NewLh + k * blk, lh + k * blk, blk * sizeof(double));
        }
    }
}

void orc_get_limiter(const orc_ctx *c, int *status, int *recompute, double *lh, double *lmin, double *lmax)
{
    size_t blk = (size_t)c->V * c->s * c->s;
    for (size_t k = 0; k < ncell(c); k++) {
        status[k] = c->status[k];
        recompute[k] = c->recompute[k];
        if (c->status[k])
            memcpy(lh + k * blk, c->Lh + k * blk, blk * sizeof(double));
        else
            memset(lh + k * blk, 0, blk * sizeof(double));
    }
    memcpy(lmin, c->lmin, c->V * ncell(c) * sizeof(double));
    memcpy(lmax, c->lmax, c->V * ncell(c) * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* per-cell routines                                                    */
/* ------------------------------------------------------------------ */

void orc_flux(orc_ctx *c, const double *q, double *f) { pde_flux(c, q, f); }
void orc_eigen(orc_ctx *c, const double *q, const double *nv, double *lam) { pde_eigen(c, q, nv, lam); }
void orc_rusanov_bn(orc_ctx *c, const double *ql, const double *fl, const double *qr, const double *fr,
                    const double *nv, double *dm, double *dp)
{
    bn_rusanov(c->eos, ql, fl, qr, fr, nv, dm, dp);
}
void orc_ncp(orc_ctx *c, const double *q, const double *gradq, double *src) { bn_fused_src_ncp(c->eos, q, gradq, src); }
void orc_roe(orc_ctx *c, const double *qa, const double *qb, const double *nv, double *b)
{
    bn_roe_matrix(c->eos, qa, qb, nv, b);
}
void orc_prim2cons(orc_ctx *c, const double *v, double *q)
{
    if (c->pde == ORC_PDE_BN)
        bn_prim2cons(c->eos, v, q);
    else
        euler_prim2cons(c->gamma, v, q);
}
void orc_rusanov(orc_ctx *c, const double *ql, const double *fl, const double *qr, const double *fr,
                 const double *nv, double *flux)
{
    euler_rusanov(c->gamma, ql, fl, qr, fr, nv, flux);
}

/* ader_space_time_predictor: Euler/space_time_predictor.f90:8-123; the Baer-Nunziato
 * variant (Baer-Nunziato/space_time_predictor.f90:8-143) adds the gradients lqx, lqy
 * (:72,79), the NCP source (:84-92) and its time average lShi (:111). */
static void predictor_impl(orc_ctx *c, double *lqhi, double *lFhi, double *lShi, double *lqbnd, double *lFbnd,
                           const double *luh)
{
    const int bn = c->pde == ORC_PDE_BN;
    static double lqx[VMAX * NMAX * NMAX * NMAX], lqy[VMAX * NMAX * NMAX * NMAX], lSh[VMAX * NMAX * NMAX * NMAX];
    const int n = c->n, V = c->V;
    const double *w = c->w;
    double lqh[VMAX * NMAX * NMAX * NMAX], rhs0[VMAX * NMAX * NMAX * NMAX], rhs[VMAX * NMAX * NMAX * NMAX];
    double lFh[2 * VMAX * NMAX * NMAX * NMAX];
#define Q4(v, i, j, l) ((v) + V * ((i) + n * ((j) + n * (l))))
#define F5(v, d, i, j, l) ((v) + V * ((d) + 2 * ((i) + n * ((j) + n * (l)))))
    /* :29-45 initial guess and rhs0 */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            for (int v = 0; v < V; v++)
                for (int l = 0; l < n; l++) lqh[Q4(v, i, j, l)] = luh[v + V * (i + n * j)];
            for (int v = 0; v < V; v++)
                for (int l = 0; l < n; l++) rhs0[Q4(v, i, j, l)] = w[i] * w[j] * c->f0[l] * luh[v + V * (i + n * j)];
        }
    OPS(3 * V * n * n * n);
    /* :49-84 Picard iterations */
    for (int iter = 0; iter < n; iter++) {
        for (int l = 0; l < n; l++) {
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) pde_flux(c, &lqh[Q4(0, i, j, l)], &lFh[F5(0, 0, i, j, l)]);
            for (int j = 0; j < n; j++) { /* :65-67 */
                double sx = w[l] * w[j] * c->dt / c->dx[0];
                for (int i = 0; i < n; i++)
                    for (int v = 0; v < V; v++) {
                        double acc = 0.0;
                        for (int k = 0; k < n; k++) acc = acc + lFh[F5(v, 0, k, j, l)] * c->kxi[k + n * i];
                        rhs[Q4(v, i, j, l)] = rhs0[Q4(v, i, j, l)] - sx * acc;
                        if (bn) { /* Baer-Nunziato/space_time_predictor.f90:72 */
                            acc = 0.0;
                            for (int k = 0; k < n; k++) acc = acc + lqh[Q4(v, k, j, l)] * c->dudx[i + n * k];
                            lqx[Q4(v, i, j, l)] = 1.0 / c->dx[0] * acc;
                        }
                    }
            }
            for (int i = 0; i < n; i++) { /* :71-73 */
                double sy = w[l] * w[i] * c->dt / c->dx[1];
                for (int j = 0; j < n; j++)
                    for (int v = 0; v < V; v++) {
                        double acc = 0.0;
                        for (int k = 0; k < n; k++) acc = acc + lFh[F5(v, 1, i, k, l)] * c->kxi[k + n * j];
                        rhs[Q4(v, i, j, l)] = rhs[Q4(v, i, j, l)] - sy * acc;
                        if (bn) { /* :79 */
                            acc = 0.0;
                            for (int k = 0; k < n; k++) acc = acc + lqh[Q4(v, i, k, l)] * c->dudx[j + n * k];
                            lqy[Q4(v, i, j, l)] = 1.0 / c->dx[1] * acc;
                        }
                    }
            }
            if (bn) { /* NCP source, Baer-Nunziato/space_time_predictor.f90:84-92 */
                for (int j = 0; j < n; j++)
                    for (int i = 0; i < n; i++) {
                        double grad[18], src[9];
                        for (int v = 0; v < V; v++) {
                            grad[v] = lqx[Q4(v, i, j, l)];
                            grad[9 + v] = lqy[Q4(v, i, j, l)];
                        }
                        bn_fused_src_ncp(c->eos, &lqh[Q4(0, i, j, l)], grad, src);
                        const double f = w[i] * w[j] * w[l] * c->dt;
                        for (int v = 0; v < V; v++) {
                            lSh[Q4(v, i, j, l)] = src[v];
                            rhs[Q4(v, i, j, l)] = rhs[Q4(v, i, j, l)] + f * src[v];
                        }
                    }
                OPS(2 * V * n * n * (2 * n + 1) + n * n * (3 + 2 * V));
            }
            OPS(2 * (3 * n) + 2 * V * n * n * (2 * n + 2));
        }
        for (int j = 0; j < n; j++) /* :79-83 */
            for (int i = 0; i < n; i++) {
                double f = 1.0 / (w[i] * w[j]);
                double tmp[VMAX * NMAX];
                for (int l = 0; l < n; l++)
                    for (int v = 0; v < V; v++) {
                        double acc = 0.0;
                        for (int m = 0; m < n; m++) acc = acc + rhs[Q4(v, i, j, m)] * c->ik1[l + n * m];
                        tmp[v + V * l] = f * acc;
                    }
                for (int l = 0; l < n; l++)
                    for (int v = 0; v < V; v++) lqh[Q4(v, i, j, l)] = tmp[v + V * l];
            }
        OPS(n * n * (2 + V * n * (2 * n + 1)));
    }
    /* :88-95 time averages */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int l = 0; l < n; l++) acc = acc + lqh[Q4(v, i, j, l)] * w[l];
                lqhi[v + V * (i + n * j)] = acc;
            }
            if (bn)
                for (int v = 0; v < V; v++) { /* Baer-Nunziato/space_time_predictor.f90:111 */
                    double acc = 0.0;
                    for (int l = 0; l < n; l++) acc = acc + lSh[Q4(v, i, j, l)] * w[l];
                    lShi[v + V * (i + n * j)] = acc;
                }
            for (int d = 0; d < 2; d++)
                for (int v = 0; v < V; v++) {
                    double acc = 0.0;
                    for (int l = 0; l < n; l++) acc = acc + lFh[F5(v, d, i, j, l)] * w[l];
                    lFhi[v + V * (d + 2 * (i + n * j))] = acc;
                }
        }
    OPS((bn ? 4 : 3) * V * n * n * 2 * n);
    /* :100-119 boundary extrapolation */
    for (int j = 0; j < n; j++)
        for (int v = 0; v < V; v++) {
            double a = 0.0, b = 0.0, e = 0.0, g = 0.0;
            for (int i = 0; i < n; i++) {
                a = a + lqhi[v + V * (i + n * j)] * c->flc[i];
                b = b + lqhi[v + V * (i + n * j)] * c->frc[i];
                e = e + lFhi[v + V * (0 + 2 * (i + n * j))] * c->flc[i];
                g = g + lFhi[v + V * (0 + 2 * (i + n * j))] * c->frc[i];
            }
            lqbnd[v + V * (0 + 4 * j)] = a;
            lqbnd[v + V * (1 + 4 * j)] = b;
            lFbnd[v + V * (0 + 4 * j)] = e;
            lFbnd[v + V * (1 + 4 * j)] = g;
        }
    for (int i = 0; i < n; i++)
        for (int v = 0; v < V; v++) {
            double a = 0.0, b = 0.0, e = 0.0, g = 0.0;
            for (int j = 0; j < n; j++) {
                a = a + lqhi[v + V * (i + n * j)] * c->flc[j];
                b = b + lqhi[v + V * (i + n * j)] * c->frc[j];
                e = e + lFhi[v + V * (1 + 2 * (i + n * j))] * c->flc[j];
                g = g + lFhi[v + V * (1 + 2 * (i + n * j))] * c->frc[j];
            }
            lqbnd[v + V * (2 + 4 * i)] = a;
            lqbnd[v + V * (3 + 4 * i)] = b;
            lFbnd[v + V * (2 + 4 * i)] = e;
            lFbnd[v + V * (3 + 4 * i)] = g;
        }
    OPS(8 * V * n * 2 * n);
#undef Q4
#undef F5
}

void orc_predictor(orc_ctx *c, double *lqhi, double *lFhi, double *lqbnd, double *lFbnd, const double *luh)
{
    double shi[VMAX * NMAX * NMAX];
    predictor_impl(c, lqhi, lFhi, shi, lqbnd, lFbnd, luh);
}

void orc_predictor_bn(orc_ctx *c, double *lqhi, double *lFhi, double *lShi, double *lqbnd, double *lFbnd,
                      const double *luh)
{
    predictor_impl(c, lqhi, lFhi, lShi, lqbnd, lFbnd, luh);
}

/* volume_integral: Euler/volume_integral.f90:8-36; Baer-Nunziato adds w_i w_j lShi
 * (Baer-Nunziato/volume_integral.f90:36-40) when lShi is given */
static void volume_impl(orc_ctx *c, double *lduh, const double *lFhi, const double *lShi);
void orc_volume(orc_ctx *c, double *lduh, const double *lFhi) { volume_impl(c, lduh, lFhi, NULL); }
void orc_volume_bn(orc_ctx *c, double *lduh, const double *lFhi, const double *lShi) { volume_impl(c, lduh, lFhi, lShi); }
static void volume_impl(orc_ctx *c, double *lduh, const double *lFhi, const double *lShi)
{
    const int n = c->n, V = c->V;
    for (int k = 0; k < V * n * n; k++) lduh[k] = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int k = 0; k < n; k++) acc = acc + lFhi[v + V * (0 + 2 * (k + n * j))] * c->kxi[i + n * k];
                lduh[v + V * (i + n * j)] = lduh[v + V * (i + n * j)] + acc * c->w[j] / c->dx[0];
            }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int k = 0; k < n; k++) acc = acc + lFhi[v + V * (1 + 2 * (i + n * k))] * c->kxi[j + n * k];
                lduh[v + V * (i + n * j)] = lduh[v + V * (i + n * j)] + acc * c->w[i] / c->dx[1];
            }
    if (lShi)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                for (int v = 0; v < V; v++)
                    lduh[v + V * (i + n * j)] = lduh[v + V * (i + n * j)] + c->w[i] * c->w[j] * lShi[v + V * (i + n * j)];
    OPS(2 * V * n * n * (2 * n + 3) + (lShi ? 3 * V * n * n : 0));
}

/* surface_integral: Euler/surface_integral.f90:8-39 */
void orc_surface(orc_ctx *c, double *lduh, const double *fl, const double *fr, const double *gb, const double *gt)
{
    const int n = c->n, V = c->V;
    for (int j = 0; j < n; j++)
        for (int v = 0; v < V; v++) {
            double s = c->w[j] / c->dx[0];
            for (int i = 0; i < n; i++)
                lduh[v + V * (i + n * j)] =
                    lduh[v + V * (i + n * j)] - s * (fr[v + V * j] * c->frc[i] - fl[v + V * j] * c->flc[i]);
        }
    for (int i = 0; i < n; i++)
        for (int v = 0; v < V; v++) {
            double s = c->w[i] / c->dx[1];
            for (int j = 0; j < n; j++)
                lduh[v + V * (i + n * j)] =
                    lduh[v + V * (i + n * j)] - s * (gt[v + V * i] * c->frc[j] - gb[v + V * i] * c->flc[j]);
        }
    OPS(2 * V * n * (1 + 5 * n));
}

/* element_update: Euler/element_update.f90:8-30 */
void orc_update(orc_ctx *c, double *luh, double *lduh)
{
    const int n = c->n, V = c->V;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            for (int v = 0; v < V; v++) lduh[v + V * (i + n * j)] = lduh[v + V * (i + n * j)] / (c->w[i] * c->w[j]);
    for (int k = 0; k < V * n * n; k++) luh[k] = luh[k] + c->dt * lduh[k];
    OPS(4 * V * n * n);
}

/* get_subcell_data: Euler/limiter.f90:12-37 */
void orc_subcell_data(orc_ctx *c, double *lim, const double *luh)
{
    const int n = c->n, V = c->V, s = c->s;
    double limy[VMAX * NMAX * SMAX]; /* (V,n,s) */
    for (int a = 0; a < n; a++)
        for (int q = 0; q < s; q++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int b = 0; b < n; b++) acc = acc + luh[v + V * (a + n * b)] * c->u2l[q + s * b];
                limy[v + V * (a + n * q)] = acc;
            }
    for (int q = 0; q < s; q++)
        for (int p = 0; p < s; p++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int a = 0; a < n; a++) acc = acc + limy[v + V * (a + n * q)] * c->u2l[p + s * a];
                lim[v + V * (p + s * q)] = acc;
            }
    OPS(2 * V * n * s * (n + s));
}

/* get_lobatto_data: Euler/limiter.f90:43-68 */
void orc_lobatto_data(orc_ctx *c, double *lob, const double *luh)
{
    const int n = c->n, V = c->V;
    double loby[VMAX * NMAX * NMAX];
    for (int a = 0; a < n; a++)
        for (int q = 0; q < n; q++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int b = 0; b < n; b++) acc = acc + luh[v + V * (a + n * b)] * c->u2b[q + n * b];
                loby[v + V * (a + n * q)] = acc;
            }
    for (int q = 0; q < n; q++)
        for (int p = 0; p < n; p++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int a = 0; a < n; a++) acc = acc + loby[v + V * (a + n * q)] * c->u2b[p + n * a];
                lob[v + V * (p + n * q)] = acc;
            }
    OPS(4 * V * n * n * n);
}

/* put_subcell_data: Euler/limiter.f90:74-98 */
void orc_put_subcell_data(orc_ctx *c, double *luh, const double *lim)
{
    const int n = c->n, V = c->V, s = c->s;
    double luhy[VMAX * SMAX * NMAX]; /* (V,s,n) */
    for (int p = 0; p < s; p++)
        for (int b = 0; b < n; b++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int q = 0; q < s; q++) acc = acc + lim[v + V * (p + s * q)] * c->l2u[b + n * q];
                luhy[v + V * (p + s * b)] = acc;
            }
    for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++)
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                for (int p = 0; p < s; p++) acc = acc + luhy[v + V * (p + s * b)] * c->l2u[a + n * p];
                luh[v + V * (a + n * b)] = acc;
            }
    OPS(2 * V * n * s * (s + n));
}

/* min/max over nodal, subcell and Lobatto values of one variable.
 * dmp_order = 0: min(minval(uh), minval(lob), minval(lim))   (get_min_max, Euler/limiter.f90:125-131)
 * dmp_order = 1: min(minval(uh), minval(lim), minval(lob))   (DMP, Euler/limiter.f90:202-208) */
static void cell_range(orc_ctx *c, const double *luh, const double *lim, const double *lob, int v, int dmp_order,
                       double *mn, double *mx)
{
    const int n = c->n, V = c->V, s = c->s;
    double u0 = minval_s(luh + v, n * n, V), u1 = maxval_s(luh + v, n * n, V);
    double b0 = minval_s(lob + v, n * n, V), b1 = maxval_s(lob + v, n * n, V);
    double l0 = minval_s(lim + v, s * s, V), l1 = maxval_s(lim + v, s * s, V);
    if (dmp_order) {
        *mn = f_min(f_min(u0, l0), b0);
        *mx = f_max(f_max(u1, l1), b1);
    } else {
        *mn = f_min(f_min(u0, b0), l0);
        *mx = f_max(f_max(u1, b1), l1);
    }
}

/* DMP: Euler/limiter.f90:179-245; returns 1 if the cell is admissible */
int orc_dmp(orc_ctx *c, const double *luh, const double *lmin, const double *lmax, double argmax)
{
    const int V = c->V, s = c->s;
    double lim[VMAX * SMAX * SMAX], lob[VMAX * NMAX * NMAX];
    orc_subcell_data(c, lim, luh);
    orc_lobatto_data(c, lob, luh);
    for (int v = 0; v < V; v++) {
        double mn, mx;
        cell_range(c, luh, lim, lob, v, 1, &mn, &mx);
        double ldiff = f_max(f_max(1.0e-4, 1.0e-3 * (lmax[v] - lmin[v])), argmax);
        OPS(2 + 2);
        if (mn < lmin[v] - ldiff) return 0;
        if (mx > lmax[v] + ldiff) return 0;
    }
    for (int k = 0; k < s * s; k++)
        if (pde_positivity(c, &lim[V * k]) < 0) return 0;
    return 1;
}

double orc_dmp_margin(orc_ctx *c, const double *luh, const double *lmin, const double *lmax, double argmax)
{
    const int V = c->V, s = c->s;
    double lim[VMAX * SMAX * SMAX], lob[VMAX * NMAX * NMAX];
    orc_subcell_data(c, lim, luh);
    orc_lobatto_data(c, lob, luh);
    double margin = INFINITY;
    for (int v = 0; v < V; v++) {
        double mn, mx;
        cell_range(c, luh, lim, lob, v, 1, &mn, &mx);
        double ldiff = f_max(f_max(1.0e-4, 1.0e-3 * (lmax[v] - lmin[v])), argmax);
        double a = mn - (lmin[v] - ldiff), b = (lmax[v] + ldiff) - mx;
        if (a < margin) margin = a;
        if (b < margin) margin = b;
    }
    for (int k = 0; k < s * s && c->pde == ORC_PDE_BN; k++)
        if (bn_positivity(c->eos, &lim[V * k]) < 0 && margin > -1.0) margin = -1.0;
    for (int k = 0; k < s * s && c->pde == ORC_PDE_EULER; k++) {
        const double *Q = &lim[V * k];
        double irho = 1.0 / Q[0];
        double p = (c->gamma - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) * irho);
        if (Q[0] - 1.0e-12 < margin) margin = Q[0] - 1.0e-12;
        if (p - 1.0e-12 < margin) margin = p - 1.0e-12;
    }
    return margin;
}

/* ------------------------------------------------------------------ */
/* whole-mesh phases                                                    */
/* ------------------------------------------------------------------ */

#define UH(c, arr, ci, cj) (&(arr)[(size_t)(c)->V * (c)->n * (c)->n * ((size_t)(ci) + (size_t)(c)->imax * (cj))])
#define BND(c, arr, ci, cj) (&(arr)[(size_t)(c)->V * 4 * (c)->n * ((size_t)(ci) + (size_t)(c)->imax * (cj))])

/* neighbour table of Euler/initialize.f90:226-258: clamp, wrap along an axis whose
 * sides are periodic */
static void neighbor(const orc_ctx *c, int i, int j, int di, int dj, int *in, int *jn)
{
    int a = i + di, b = j + dj;
    const int dmr = c->bc_mode == ORC_BC_DMR;
    const int px = !dmr && c->side[0] == 3, py = !dmr && c->side[2] == 3;
    if (a < 0) a = px ? c->imax - 1 : 0;
    if (a > c->imax - 1) a = px ? 0 : c->imax - 1;
    if (b < 0) b = py ? c->jmax - 1 : 0;
    if (b > c->jmax - 1) b = py ? 0 : c->jmax - 1;
    *in = a;
    *jn = b;
}

void orc_set_sides(orc_ctx *c, const int codes[4], const double *qin)
{
    c->bc_mode = ORC_BC_SIDES;
    for (int k = 0; k < 4; k++) c->side[k] = codes[k];
    for (int v = 0; v < c->V; v++) c->qin[v] = qin ? qin[v] : 0.0;
}

/* the momentum component of direction d (0 x, 1 y) of either phase */
static int normal_momentum(const orc_ctx *c, int v, int d)
{
    return v == 1 + d || (c->V == 9 && v == 5 + d);
}

/* Ghost state and normal flux of a non-periodic side (orc_set_sides): inflow takes qin
 * and its flux along nv, computed as the DMR inflow (Euler/riemann_problem.f90:186-193);
 * transmissive copies the inside face (Baer-Nunziato/riemann_problem.f90:26-34);
 * reflective copies it with the normal momentum negated and the other flux
 * components negated, as the DMR wall (Euler/riemann_problem.f90:225-228). */
static void side_ghost(const orc_ctx *c, int code, const double *nv, const double *qi, const double *fi, double *Q,
                       double *F)
{
    const int V = c->V;
    if (code == 0) {
        double Fbc[2 * VMAX];
        for (int v = 0; v < V; v++) Q[v] = c->qin[v];
        pde_flux(c, Q, Fbc);
        for (int v = 0; v < V; v++) F[v] = 0.0 + Fbc[v] * nv[0] + Fbc[v + V] * nv[1];
        return;
    }
    for (int v = 0; v < V; v++) {
        Q[v] = qi[v];
        F[v] = fi[v];
    }
    if (code == 2) {
        const int d = nv[0] != 0.0 ? 0 : 1;
        for (int v = 0; v < V; v++) {
            if (normal_momentum(c, v, d))
                Q[v] = -Q[v];
            else
                F[v] = -F[v];
        }
    }
}

/* Ghost subcells of a boundary cell (ci, cj) in the (3s)^2 block array addressed by
 * at(v, ii, jj), ii, jj in 1-s..2s: inflow replaces, reflective mirrors the normal
 * momentum; left and right, then bottom and top (the DMR overrides of
 * Euler/limiter.f90:491-522 for a code per side). */
static void side_subcells(const orc_ctx *c, int ci, int cj, double *subuh0)
{
    const int V = c->V, s = c->s, W = 3 * s;
#define AT(v, ii, jj) subuh0[(v) + V * (((ii) + s - 1) + W * ((jj) + s - 1))]
    for (int sd = 0; sd < 4; sd++) {
        const int code = c->side[sd];
        if (code != 0 && code != 2) continue;
        if ((sd == 0 && ci != 0) || (sd == 1 && ci != c->imax - 1) || (sd == 2 && cj != 0) ||
            (sd == 3 && cj != c->jmax - 1))
            continue;
        const int d = sd < 2 ? 0 : 1;
        const int lo = (sd == 0 || sd == 2) ? 1 - s : s + 1, hi = (sd == 0 || sd == 2) ? 0 : 2 * s;
        for (int b = 1 - s; b <= 2 * s; b++)
            for (int a = lo; a <= hi; a++)
                for (int v = 0; v < V; v++) {
                    double *x = d == 0 ? &AT(v, a, b) : &AT(v, b, a);
                    if (code == 0)
                        *x = c->qin[v];
                    else if (normal_momentum(c, v, d))
                        *x = -*x;
                }
    }
#undef AT
}

/* calc_time_step: Euler/calc_time_step.f90:9-52 */
static double calc_dt_ad(orc_ctx *c);
double orc_calc_dt(orc_ctx *c)
{
    if (c->pde == ORC_PDE_AD) return calc_dt_ad(c);
    const int n = c->n, V = c->V;
    const double nvx[2] = {1.0, 0.0}, nvy[2] = {0.0, 1.0};
    double dt = 1.0e20;
    for (int cj = 0; cj < c->jmax; cj++)
        for (int ci = 0; ci < c->imax; ci++) {
            const double *u = UH(c, c->uh, ci, cj);
            for (int k = 0; k < n; k++)
                for (int l = 0; l < n; l++) {
                    double L[VMAX], denom = 0.0;
                    pde_eigen(c, &u[V * (k + n * l)], nvx, L);
                    denom = denom + maxabs(L, V) / c->dx[0];
                    pde_eigen(c, &u[V * (k + n * l)], nvy, L);
                    denom = denom + maxabs(L, V) / c->dx[1];
                    dt = f_min(dt, c->cflpnpm / denom);
                    OPS(5);
                }
        }
    if (c->time + dt > c->tend) dt = c->tend - c->time;
    c->dt = dt;
    return dt;
}

/* get_min_max: Euler/limiter.f90:104-171 */
static void get_min_max(orc_ctx *c)
{
    const int V = c->V, s = c->s;
    double lim[VMAX * SMAX * SMAX], lob[VMAX * NMAX * NMAX];
    for (int cj = 0; cj < c->jmax; cj++)
        for (int ci = 0; ci < c->imax; ci++) {
            size_t k = (size_t)ci + (size_t)c->imax * cj;
            if (c->status[k] == 0) {
                const double *u = UH(c, c->uh, ci, cj);
                orc_subcell_data(c, lim, u);
                orc_lobatto_data(c, lob, u);
                for (int v = 0; v < V; v++)
                    cell_range(c, u, lim, lob, v, 0, &c->lmin_old[V * k + v], &c->lmax_old[V * k + v]);
            } else {
                const double *L = &c->Lh[(size_t)V * s * s * k];
                for (int v = 0; v < V; v++) {
                    c->lmin_old[V * k + v] = minval_s(L + v, s * s, V);
                    c->lmax_old[V * k + v] = maxval_s(L + v, s * s, V);
                }
            }
        }
    for (int cj = 0; cj < c->jmax; cj++)
        for (int ci = 0; ci < c->imax; ci++) {
            size_t k = (size_t)ci + (size_t)c->imax * cj;
            double qmn[VMAX], qmx[VMAX];
            for (int v = 0; v < V; v++) {
                qmn[v] = c->lmin_old[V * k + v];
                qmx[v] = c->lmax_old[V * k + v];
            }
            for (int jj = -1; jj <= 1; jj++)
                for (int ii = -1; ii <= 1; ii++) {
                    int in, jn;
                    neighbor(c, ci, cj, ii, jj, &in, &jn);
                    size_t kn = (size_t)in + (size_t)c->imax * jn;
                    for (int v = 0; v < V; v++) {
                        qmn[v] = f_min(qmn[v], c->lmin_old[V * kn + v]);
                        qmx[v] = f_max(qmx[v], c->lmax_old[V * kn + v]);
                    }
                }
            for (int v = 0; v < V; v++) {
                c->lmin[V * k + v] = qmn[v];
                c->lmax[V * k + v] = qmx[v];
            }
        }
}

/* DMR boundary states of Euler/riemann_problem.f90:165-175 and the shock position of :1-8 */
static void dmr_states(const orc_ctx *c, double *Qa, double *Qb, double *xsw)
{
    const double PI = 4.0 * atan(1.0);
    double Va[4] = {8.0, 8.25 * cos(PI / 6.0), -8.25 * sin(PI / 6.0), 116.5};
    double Vb[4] = {1.4, 0.0, 0.0, 1.0};
    euler_prim2cons(c->gamma, Va, Qa);
    euler_prim2cons(c->gamma, Vb, Qb);
    *xsw = 1.0 / 6.0 + 1.0 / tan(PI / 3.0) + (10.0 / sin(PI / 3.0)) * c->time;
}

/* solve_riemann_problem: Euler/riemann_problem.f90:10-116 (DMR boundaries), or the
 * boundary states of the side codes (side_ghost; interior faces identical) */
static void riemann(orc_ctx *c)
{
    const int n = c->n, V = c->V, I = c->imax, J = c->jmax;
    double QL[4], QR[4], FL[4], FR[4], Qa[4], Qb[4], Fbc[8], xsw;
    dmr_states(c, Qa, Qb, &xsw);
    double nv[2] = {1.0, 0.0};
    for (int j = 0; j < J; j++)
        for (int i = 0; i <= I; i++)
            for (int l = 0; l < n; l++) {
                if (i == 0) {
                    const double *qr = &BND(c, c->qbnd, 0, j)[V * (0 + 4 * l)];
                    const double *fr = &BND(c, c->Fbnd, 0, j)[V * (0 + 4 * l)];
                    if (c->bc_mode == ORC_BC_DMR) {
                        euler_flux(c->gamma, Qa, Fbc);
                        for (int v = 0; v < V; v++) {
                            QL[v] = Qa[v];
                            FL[v] = 0.0 + Fbc[v] * nv[0] + Fbc[v + V] * nv[1];
                        }
                    } else if (c->side[0] == 3) {
                        for (int v = 0; v < V; v++) {
                            QL[v] = BND(c, c->qbnd, I - 1, j)[v + V * (1 + 4 * l)];
                            FL[v] = BND(c, c->Fbnd, I - 1, j)[v + V * (1 + 4 * l)];
                        }
                    } else {
                        side_ghost(c, c->side[0], nv, qr, fr, QL, FL);
                    }
                    for (int v = 0; v < V; v++) {
                        QR[v] = qr[v];
                        FR[v] = fr[v];
                    }
                } else if (i == I) {
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, I - 1, j)[v + V * (1 + 4 * l)];
                        FL[v] = BND(c, c->Fbnd, I - 1, j)[v + V * (1 + 4 * l)];
                    }
                    if (c->bc_mode != ORC_BC_DMR && c->side[1] == 3) {
                        for (int v = 0; v < V; v++) {
                            QR[v] = BND(c, c->qbnd, 0, j)[v + V * (0 + 4 * l)];
                            FR[v] = BND(c, c->Fbnd, 0, j)[v + V * (0 + 4 * l)];
                        }
                    } else {  /* the DMR right side is transmissive (:195-198) */
                        side_ghost(c, c->bc_mode == ORC_BC_DMR ? 1 : c->side[1], nv, QL, FL, QR, FR);
                    }
                } else {
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, i - 1, j)[v + V * (1 + 4 * l)];
                        FL[v] = BND(c, c->Fbnd, i - 1, j)[v + V * (1 + 4 * l)];
                        QR[v] = BND(c, c->qbnd, i, j)[v + V * (0 + 4 * l)];
                        FR[v] = BND(c, c->Fbnd, i, j)[v + V * (0 + 4 * l)];
                    }
                }
                euler_rusanov(c->gamma, QL, FL, QR, FR, nv, &c->Fu[V * (l + n * (i + (I + 1) * j))]);
            }
    nv[0] = 0.0;
    nv[1] = 1.0;
    for (int j = 0; j <= J; j++)
        for (int i = 0; i < I; i++)
            for (int k = 0; k < n; k++) {
                if (j == 0) {
                    const double *qr = &BND(c, c->qbnd, i, 0)[V * (2 + 4 * k)];
                    const double *fr = &BND(c, c->Fbnd, i, 0)[V * (2 + 4 * k)];
                    if (c->bc_mode != ORC_BC_DMR && c->side[2] == 3) {
                        for (int v = 0; v < V; v++) {
                            QL[v] = BND(c, c->qbnd, i, J - 1)[v + V * (3 + 4 * k)];
                            FL[v] = BND(c, c->Fbnd, i, J - 1)[v + V * (3 + 4 * k)];
                        }
                    } else if (c->bc_mode != ORC_BC_DMR) {
                        side_ghost(c, c->side[2], nv, qr, fr, QL, FL);
                    } else {
                        for (int v = 0; v < V; v++) {
                            QL[v] = qr[v];
                            FL[v] = fr[v];
                        }
                        if (!((double)(i + 1) * c->dx[0] <= 1.0 / 6.0)) {
                            QL[2] = -QL[2];
                            FL[0] = -FL[0];
                            FL[1] = -FL[1];
                            FL[3] = -FL[3];
                        }
                    }
                    for (int v = 0; v < V; v++) {
                        QR[v] = qr[v];
                        FR[v] = fr[v];
                    }
                } else if (j == J) {
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, i, J - 1)[v + V * (3 + 4 * k)];
                        FL[v] = BND(c, c->Fbnd, i, J - 1)[v + V * (3 + 4 * k)];
                    }
                    if (c->bc_mode == ORC_BC_DMR) {
                        const double *Qbc = ((double)(i + 1) * c->dx[0] <= xsw) ? Qa : Qb;
                        euler_flux(c->gamma, Qbc, Fbc);
                        for (int v = 0; v < V; v++) {
                            QR[v] = Qbc[v];
                            FR[v] = 0.0 + Fbc[v] * nv[0] + Fbc[v + V] * nv[1];
                        }
                    } else if (c->side[3] == 3) {
                        for (int v = 0; v < V; v++) {
                            QR[v] = BND(c, c->qbnd, i, 0)[v + V * (2 + 4 * k)];
                            FR[v] = BND(c, c->Fbnd, i, 0)[v + V * (2 + 4 * k)];
                        }
                    } else {
                        side_ghost(c, c->side[3], nv, QL, FL, QR, FR);
                    }
                } else {
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, i, j - 1)[v + V * (3 + 4 * k)];
                        FL[v] = BND(c, c->Fbnd, i, j - 1)[v + V * (3 + 4 * k)];
                        QR[v] = BND(c, c->qbnd, i, j)[v + V * (2 + 4 * k)];
                        FR[v] = BND(c, c->Fbnd, i, j)[v + V * (2 + 4 * k)];
                    }
                }
                euler_rusanov(c->gamma, QL, FL, QR, FR, nv, &c->Gu[V * (k + n * (i + I * j))]);
            }
}

void orc_get_face_fluxes(const orc_ctx *c, double *fu, double *gu)
{
    memcpy(fu, c->Fu, sizeof(double) * c->V * c->n * (c->imax + 1) * c->jmax);
    memcpy(gu, c->Gu, sizeof(double) * c->V * c->n * c->imax * (c->jmax + 1));
}

static void minmod(double *out, const double *a, const double *b, int V)
{
    for (int v = 0; v < V; v++) {
        if (a[v] * b[v] <= 0.0)
            out[v] = 0.0;
        else
            out[v] = fabs(a[v]) < fabs(b[v]) ? a[v] : b[v];
    }
}

/* RusanovFluxS: Euler/limiter.f90:361-378 */
static void rusanov_s(double g, double *flux, const double *QL, const double *QR, const double *nv)
{
    double LL[4], LR[4], FL[8], FR[8];
    euler_eigen(g, QL, nv, LL);
    euler_eigen(g, QR, nv, LR);
    double smax = f_max(maxabs(LL, 4), maxabs(LR, 4));
    euler_flux(g, QL, FL);
    euler_flux(g, QR, FR);
    for (int v = 0; v < 4; v++) {
        double m = 0.0;
        m = m + (FL[v] + FR[v]) * nv[0];
        m = m + (FL[v + 4] + FR[v + 4]) * nv[1];
        flux[v] = 0.5 * m - 0.5 * smax * (QR[v] - QL[v]);
    }
    OPS(4 * 11);
}

/* subcell_recompute: Euler/limiter.f90:411-647 */
static void subcell_recompute(orc_ctx *c)
{
    const int V = c->V, s = c->s, I = c->imax, J = c->jmax;
    const int W = 3 * s;             /* subuh0 index range 1-s .. 2s -> 0 .. 3s-1 */
    const int H = s + 2;             /* slopes/w range 0 .. s+1 (dn = 1) */
    double Qa[4], Qb[4], xsw;
    dmr_states(c, Qa, Qb, &xsw);
    double ldx[2] = {c->dx[0] / (double)s, c->dx[1] / (double)s};
    const double nvx[2] = {1.0, 0.0}, nvy[2] = {0.0, 1.0};
    double *subuh0 = (double *)malloc(sizeof(double) * V * W * W);
    double *slx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *sly = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wLx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wRx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wLy = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wRy = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *Fx = (double *)calloc((size_t)V * (s + 1) * s, sizeof(double));
    double *Fy = (double *)calloc((size_t)V * s * (s + 1), sizeof(double));
    double lim[VMAX * SMAX * SMAX], sub1[VMAX * SMAX * SMAX];
#define S0(v, ii, jj) subuh0[(v) + V * (((ii) + s - 1) + W * ((jj) + s - 1))] /* ii,jj in 1-s..2s */
#define HW(a, v, ii, jj) (a)[(v) + V * ((ii) + H * (jj))]                     /* ii,jj in 0..s+1 */
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            size_t k = (size_t)ci + (size_t)I * cj;
            if (c->recompute[k] <= 0) continue;
            for (int jj = -1; jj <= 1; jj++)
                for (int ii = -1; ii <= 1; ii++) {
                    int in, jn;
                    neighbor(c, ci, cj, ii, jj, &in, &jn);
                    size_t kn = (size_t)in + (size_t)I * jn;
                    const double *src;
                    if (c->oldstatus[kn] == 0) {
                        orc_subcell_data(c, lim, UH(c, c->olduh, in, jn));
                        src = lim;
                    } else {
                        src = &c->Lh[(size_t)V * s * s * kn];
                    }
                    for (int q = 1; q <= s; q++)
                        for (int p = 1; p <= s; p++)
                            for (int v = 0; v < V; v++) S0(v, p + ii * s, q + jj * s) = src[v + V * ((p - 1) + s * (q - 1))];
                }
            if (c->bc_mode != ORC_BC_DMR) side_subcells(c, ci, cj, subuh0);
            if (c->bc_mode == ORC_BC_DMR) { /* :491-522 */
                if (ci == 0)
                    for (int jj = 1 - s; jj <= 2 * s; jj++)
                        for (int ii = 1 - s; ii <= 0; ii++)
                            for (int v = 0; v < V; v++) S0(v, ii, jj) = Qa[v];
                if (cj == 0 && (double)(ci + 1) * c->dx[0] > 1.0 / 6.0)
                    for (int jj = 1 - s; jj <= 0; jj++)
                        for (int ii = 1 - s; ii <= 2 * s; ii++) S0(2, ii, jj) = -S0(2, ii, jj);
                if (cj == J - 1) {
                    const double *Qbc = ((double)(ci + 1) * c->dx[0] <= xsw) ? Qa : Qb;
                    for (int jj = 1 + s; jj <= 2 * s; jj++)
                        for (int ii = 1 - s; ii <= 2 * s; ii++)
                            for (int v = 0; v < V; v++) S0(v, ii, jj) = Qbc[v];
                }
            }
            /* :529-536 minmod slopes on the interior subcells only */
            for (int jj = 1; jj <= s; jj++)
                for (int ii = 1; ii <= s; ii++) {
                    double a[VMAX], b[VMAX];
                    for (int v = 0; v < V; v++) {
                        a[v] = S0(v, ii + 1, jj) - S0(v, ii, jj);
                        b[v] = S0(v, ii, jj) - S0(v, ii - 1, jj);
                    }
                    minmod(&HW(slx, 0, ii, jj), a, b, V);
                    for (int v = 0; v < V; v++) {
                        a[v] = S0(v, ii, jj + 1) - S0(v, ii, jj);
                        b[v] = S0(v, ii, jj) - S0(v, ii, jj - 1);
                    }
                    minmod(&HW(sly, 0, ii, jj), a, b, V);
                    OPS(4 * V);
                }
            /* :540-569 half-step prediction on interior plus one halo layer */
            for (int jj = 0; jj <= s + 1; jj++)
                for (int ii = 0; ii <= s + 1; ii++) {
                    double FLx[8], FRx[8], FLy[8], FRy[8], Qt[VMAX];
                    for (int v = 0; v < V; v++) {
                        HW(wLx, v, ii, jj) = S0(v, ii, jj) - 0.5 * HW(slx, v, ii, jj);
                        HW(wRx, v, ii, jj) = S0(v, ii, jj) + 0.5 * HW(slx, v, ii, jj);
                    }
                    euler_flux(c->gamma, &HW(wLx, 0, ii, jj), FLx);
                    euler_flux(c->gamma, &HW(wRx, 0, ii, jj), FRx);
                    for (int v = 0; v < V; v++) Qt[v] = -(FRx[v] - FLx[v]) / ldx[0];
                    for (int v = 0; v < V; v++) {
                        HW(wLy, v, ii, jj) = S0(v, ii, jj) - 0.5 * HW(sly, v, ii, jj);
                        HW(wRy, v, ii, jj) = S0(v, ii, jj) + 0.5 * HW(sly, v, ii, jj);
                    }
                    euler_flux(c->gamma, &HW(wLy, 0, ii, jj), FLy);
                    euler_flux(c->gamma, &HW(wRy, 0, ii, jj), FRy);
                    for (int v = 0; v < V; v++) Qt[v] = Qt[v] - (FRy[v + 4] - FLy[v + 4]) / ldx[1];
                    for (int v = 0; v < V; v++) {
                        HW(wLx, v, ii, jj) = HW(wLx, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wRx, v, ii, jj) = HW(wRx, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wLy, v, ii, jj) = HW(wLy, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wRy, v, ii, jj) = HW(wRy, v, ii, jj) + 0.5 * c->dt * Qt[v];
                    }
                    OPS(V * (4 + 5 + 12));
                }
            /* :571-588 subcell Rusanov fluxes */
            for (int jj = 1; jj <= s; jj++)
                for (int ii = 1; ii <= s + 1; ii++)
                    rusanov_s(c->gamma, &Fx[V * ((ii - 1) + (s + 1) * (jj - 1))], &HW(wRx, 0, ii - 1, jj),
                              &HW(wLx, 0, ii, jj), nvx);
            for (int jj = 1; jj <= s + 1; jj++)
                for (int ii = 1; ii <= s; ii++)
                    rusanov_s(c->gamma, &Fy[V * ((ii - 1) + s * (jj - 1))], &HW(wRy, 0, ii, jj - 1),
                              &HW(wLy, 0, ii, jj), nvy);
            /* :592-597 finite-volume update */
            for (int jj = 1; jj <= s; jj++)
                for (int ii = 1; ii <= s; ii++)
                    for (int v = 0; v < V; v++)
                        sub1[v + V * ((ii - 1) + s * (jj - 1))] =
                            S0(v, ii, jj) -
                            c->dt / ldx[0] * (Fx[v + V * (ii + (s + 1) * (jj - 1))] - Fx[v + V * ((ii - 1) + (s + 1) * (jj - 1))]) -
                            c->dt / ldx[1] * (Fy[v + V * ((ii - 1) + s * jj)] - Fy[v + V * ((ii - 1) + s * (jj - 1))]);
            OPS(V * s * s * 8);
            double *u = UH(c, c->uh, ci, cj);
            double *nl = &c->NewLh[(size_t)V * s * s * k];
            if (c->recompute[k] == 1) { /* :599-611 */
                c->status[k] = 1;
                memcpy(nl, sub1, sizeof(double) * V * s * s);
                orc_put_subcell_data(c, u, nl);
            } else if (c->recompute[k] == 2) { /* :613-638 */
                orc_put_subcell_data(c, u, sub1);
                if (orc_dmp(c, u, &c->lmin[V * k], &c->lmax[V * k], 0.0)) {
                    c->status[k] = 0;
                } else {
                    c->status[k] = 1;
                    memcpy(nl, sub1, sizeof(double) * V * s * s);
                }
            }
        }
#undef S0
#undef HW
    free(subuh0); free(slx); free(sly); free(wLx); free(wRx); free(wLy); free(wRy); free(Fx); free(Fy);
}


/* solve_riemann_problem of Baer-Nunziato/riemann_problem.f90:9-80: transmissive on all
 * sides; D- into Fu/Gu (Fm/Gm of the reference) and D+ into Fp/Gp.  The reference
 * allocates Gm with the x-face shape (Baer-Nunziato/initialize.f90:136); the arrays here
 * have the y-face shape, which keeps every element the reference reads and writes. */
static void riemann_bn(orc_ctx *c)
{
    const int n = c->n, V = c->V, I = c->imax, J = c->jmax;
    double QL[9], QR[9], FL[9], FR[9];
    double nv[2] = {1.0, 0.0};
    /* inside face states; a boundary face takes its ghost from the side code
     * (transmissive: the copy of Baer-Nunziato/riemann_problem.f90:26-34) */
    for (int j = 0; j < J; j++)
        for (int i = 0; i <= I; i++)
            for (int l = 0; l < n; l++) {
                const int wrap = c->side[0] == 3;
                if (i > 0 || wrap)
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, i > 0 ? i - 1 : I - 1, j)[v + V * (1 + 4 * l)];
                        FL[v] = BND(c, c->Fbnd, i > 0 ? i - 1 : I - 1, j)[v + V * (1 + 4 * l)];
                    }
                if (i < I || wrap)
                    for (int v = 0; v < V; v++) {
                        QR[v] = BND(c, c->qbnd, i < I ? i : 0, j)[v + V * (0 + 4 * l)];
                        FR[v] = BND(c, c->Fbnd, i < I ? i : 0, j)[v + V * (0 + 4 * l)];
                    }
                if (i == 0 && !wrap) side_ghost(c, c->side[0], nv, QR, FR, QL, FL);
                if (i == I && !wrap) side_ghost(c, c->side[1], nv, QL, FL, QR, FR);
                const size_t o = (size_t)V * (l + n * (i + (I + 1) * j));
                bn_rusanov(c->eos, QL, FL, QR, FR, nv, &c->Fu[o], &c->Fp[o]);
            }
    nv[0] = 0.0;
    nv[1] = 1.0;
    for (int j = 0; j <= J; j++)
        for (int i = 0; i < I; i++)
            for (int k = 0; k < n; k++) {
                const int wrap = c->side[2] == 3;
                if (j > 0 || wrap)
                    for (int v = 0; v < V; v++) {
                        QL[v] = BND(c, c->qbnd, i, j > 0 ? j - 1 : J - 1)[v + V * (3 + 4 * k)];
                        FL[v] = BND(c, c->Fbnd, i, j > 0 ? j - 1 : J - 1)[v + V * (3 + 4 * k)];
                    }
                if (j < J || wrap)
                    for (int v = 0; v < V; v++) {
                        QR[v] = BND(c, c->qbnd, i, j < J ? j : 0)[v + V * (2 + 4 * k)];
                        FR[v] = BND(c, c->Fbnd, i, j < J ? j : 0)[v + V * (2 + 4 * k)];
                    }
                if (j == 0 && !wrap) side_ghost(c, c->side[2], nv, QR, FR, QL, FL);
                if (j == J && !wrap) side_ghost(c, c->side[3], nv, QL, FL, QR, FR);
                const size_t o = (size_t)V * (k + n * (i + I * j));
                bn_rusanov(c->eos, QL, FL, QR, FR, nv, &c->Gu[o], &c->Gp[o]);
            }
}

/* RusanovFluxS of Baer-Nunziato/limiter.f90:359-383: flux and Roe fluctuation */
static void bn_rusanov_s(const double *e, double *F, double *D, const double *QL, const double *QR, const double *nv)
{
    double LL[9], LR[9], FL[18], FR[18], B[81];
    bn_eigen(e, QL, nv, LL);
    bn_eigen(e, QR, nv, LR);
    double smax = f_max(maxabs(LL, 9), maxabs(LR, 9));
    bn_flux(e, QL, FL);
    bn_flux(e, QR, FR);
    for (int v = 0; v < 9; v++) {
        double m = 0.0;
        m = m + (FL[v] + FR[v]) * nv[0];
        m = m + (FL[v + 9] + FR[v + 9]) * nv[1];
        F[v] = 0.5 * m - 0.5 * smax * (QR[v] - QL[v]);
    }
    bn_roe_matrix(e, QL, QR, nv, B);
    for (int r = 0; r < 9; r++) {
        double acc = 0.0;
        for (int k = 0; k < 9; k++) acc = acc + B[r + 9 * k] * (QR[k] - QL[k]);
        D[r] = acc;
    }
    OPS(9 * 11 + 9 * 2);
}

/* subcell_recompute of Baer-Nunziato/limiter.f90:416-596: MUSCL-Hancock with the NCP
 * source and Roe fluctuations; slopes also on the halo ring (:486-492).  The neighbour
 * table's clamp gives the ghost subcells (the reference's transmissive sides), and
 * inflow / reflective side codes then override them (side_subcells). */
static void subcell_recompute_bn(orc_ctx *c)
{
    const int V = c->V, s = c->s, I = c->imax, J = c->jmax;
    const int W = 3 * s, H = s + 2;
    const double ldx[2] = {c->dx[0] / (double)s, c->dx[1] / (double)s};
    const double nvx[2] = {1.0, 0.0}, nvy[2] = {0.0, 1.0};
    double *subuh0 = (double *)malloc(sizeof(double) * V * W * W);
    double *slx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *sly = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *slt = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wLx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wRx = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wLy = (double *)calloc((size_t)V * H * H, sizeof(double));
    double *wRy = (double *)calloc((size_t)V * H * H, sizeof(double));
    double lim[VMAX * SMAX * SMAX], sub1[VMAX * SMAX * SMAX];
#define S0(v, ii, jj) subuh0[(v) + V * (((ii) + s - 1) + W * ((jj) + s - 1))]
#define HW(a, v, ii, jj) (a)[(v) + V * ((ii) + H * (jj))]
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            size_t k = (size_t)ci + (size_t)I * cj;
            if (c->recompute[k] <= 0) continue;
            for (int jj = -1; jj <= 1; jj++)
                for (int ii = -1; ii <= 1; ii++) {
                    int in, jn;
                    neighbor(c, ci, cj, ii, jj, &in, &jn);
                    size_t kn = (size_t)in + (size_t)I * jn;
                    const double *src;
                    if (c->oldstatus[kn] == 0) {
                        orc_subcell_data(c, lim, UH(c, c->olduh, in, jn));
                        src = lim;
                    } else {
                        src = &c->Lh[(size_t)V * s * s * kn];
                    }
                    for (int q = 1; q <= s; q++)
                        for (int p = 1; p <= s; p++)
                            for (int v = 0; v < V; v++) S0(v, p + ii * s, q + jj * s) = src[v + V * ((p - 1) + s * (q - 1))];
                }
            side_subcells(c, ci, cj, subuh0);
            /* :486-513 slopes, NCP source, half step on interior and halo ring */
            for (int jj = 0; jj <= s + 1; jj++)
                for (int ii = 0; ii <= s + 1; ii++) {
                    double a[VMAX], b[VMAX], grad[18], src[9], FLx[18], FRx[18], FLy[18], FRy[18], Qt[VMAX];
                    for (int v = 0; v < V; v++) {
                        a[v] = S0(v, ii + 1, jj) - S0(v, ii, jj);
                        b[v] = S0(v, ii, jj) - S0(v, ii - 1, jj);
                    }
                    minmod(&HW(slx, 0, ii, jj), a, b, V);
                    for (int v = 0; v < V; v++) {
                        a[v] = S0(v, ii, jj + 1) - S0(v, ii, jj);
                        b[v] = S0(v, ii, jj) - S0(v, ii, jj - 1);
                    }
                    minmod(&HW(sly, 0, ii, jj), a, b, V);
                    for (int v = 0; v < V; v++) {
                        grad[v] = HW(slx, v, ii, jj) / ldx[0];
                        grad[9 + v] = HW(sly, v, ii, jj) / ldx[1];
                    }
                    double q0[VMAX];
                    for (int v = 0; v < V; v++) q0[v] = S0(v, ii, jj);
                    bn_fused_src_ncp(c->eos, q0, grad, src);
                    for (int v = 0; v < V; v++) {
                        HW(wLx, v, ii, jj) = S0(v, ii, jj) - 0.5 * HW(slx, v, ii, jj);
                        HW(wRx, v, ii, jj) = S0(v, ii, jj) + 0.5 * HW(slx, v, ii, jj);
                    }
                    bn_flux(c->eos, &HW(wLx, 0, ii, jj), FLx);
                    bn_flux(c->eos, &HW(wRx, 0, ii, jj), FRx);
                    for (int v = 0; v < V; v++) Qt[v] = src[v] - (FRx[v] - FLx[v]) / ldx[0];
                    for (int v = 0; v < V; v++) {
                        HW(wLy, v, ii, jj) = S0(v, ii, jj) - 0.5 * HW(sly, v, ii, jj);
                        HW(wRy, v, ii, jj) = S0(v, ii, jj) + 0.5 * HW(sly, v, ii, jj);
                    }
                    bn_flux(c->eos, &HW(wLy, 0, ii, jj), FLy);
                    bn_flux(c->eos, &HW(wRy, 0, ii, jj), FRy);
                    for (int v = 0; v < V; v++) Qt[v] = Qt[v] - (FRy[v + 9] - FLy[v + 9]) / ldx[1];
                    for (int v = 0; v < V; v++) {
                        HW(slt, v, ii, jj) = Qt[v];
                        HW(wLx, v, ii, jj) = HW(wLx, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wRx, v, ii, jj) = HW(wRx, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wLy, v, ii, jj) = HW(wLy, v, ii, jj) + 0.5 * c->dt * Qt[v];
                        HW(wRy, v, ii, jj) = HW(wRy, v, ii, jj) + 0.5 * c->dt * Qt[v];
                    }
                    OPS(V * (4 + 2 + 4 + 5 + 12));
                }
            /* :516-537 fluxes, fluctuations and the finite-volume update */
            for (int jj = 1; jj <= s; jj++)
                for (int ii = 1; ii <= s; ii++) {
                    double grad[18], src[9], qh[VMAX], fpl[9], fmi[9], gpl[9], gmi[9], Dpx[9], Dmx[9], Dpy[9], Dmy[9];
                    for (int v = 0; v < V; v++) {
                        grad[v] = HW(slx, v, ii, jj) / ldx[0];
                        grad[9 + v] = HW(sly, v, ii, jj) / ldx[1];
                        qh[v] = S0(v, ii, jj) + 0.5 * c->dt * HW(slt, v, ii, jj);
                    }
                    bn_fused_src_ncp(c->eos, qh, grad, src);
                    bn_rusanov_s(c->eos, fpl, Dpx, &HW(wRx, 0, ii, jj), &HW(wLx, 0, ii + 1, jj), nvx);
                    bn_rusanov_s(c->eos, fmi, Dmx, &HW(wRx, 0, ii - 1, jj), &HW(wLx, 0, ii, jj), nvx);
                    bn_rusanov_s(c->eos, gpl, Dpy, &HW(wRy, 0, ii, jj), &HW(wLy, 0, ii, jj + 1), nvy);
                    bn_rusanov_s(c->eos, gmi, Dmy, &HW(wRy, 0, ii, jj - 1), &HW(wLy, 0, ii, jj), nvy);
                    for (int v = 0; v < V; v++)
                        sub1[v + V * ((ii - 1) + s * (jj - 1))] =
                            S0(v, ii, jj) - c->dt / ldx[0] * (fpl[v] - fmi[v]) - c->dt / ldx[1] * (gpl[v] - gmi[v]) -
                            c->dt / ldx[0] * 0.5 * (Dpx[v] + Dmx[v]) - c->dt / ldx[1] * 0.5 * (Dpy[v] + Dmy[v]) +
                            c->dt * src[v];
                    OPS(V * (2 + 3 + 20));
                }
            double *u = UH(c, c->uh, ci, cj);
            double *nl = &c->NewLh[(size_t)V * s * s * k];
            if (c->recompute[k] == 1) {
                c->status[k] = 1;
                memcpy(nl, sub1, sizeof(double) * V * s * s);
                orc_put_subcell_data(c, u, nl);
            } else if (c->recompute[k] == 2) {
                orc_put_subcell_data(c, u, sub1);
                if (orc_dmp(c, u, &c->lmin[V * k], &c->lmax[V * k], 0.0)) {
                    c->status[k] = 0;
                } else {
                    c->status[k] = 1;
                    memcpy(nl, sub1, sizeof(double) * V * s * s);
                }
            }
        }
#undef S0
#undef HW
    free(subuh0); free(slx); free(sly); free(slt); free(wLx); free(wRx); free(wLy); free(wRy);
}

/* ------------------------------------------------------------------ */
/* advection-diffusion (Advection-Diffusion/), nVar = 1, no limiter     */
/* ------------------------------------------------------------------ */

void orc_set_ad(orc_ctx *c, double cx, double cy, double a, const int bc[4], int fixed)
{
    c->ad_c[0] = cx;
    c->ad_c[1] = cy;
    c->ad_a = a;
    for (int k = 0; k < 4; k++) c->ad_bc[k] = bc[k];
    c->ad_fixed = fixed;
}

/* PDEViscFlux (Advection-Diffusion/pde.f90:54-64) or PDEFusedViscConvFlux (:115-129) */
static void ad_flux(const orc_ctx *c, int fused, double q, const double *grad, double *F)
{
    const double fv0 = -(c->ad_a * grad[0]), fv1 = -(c->ad_a * grad[1]);
    if (fused) {
        F[0] = c->ad_c[0] * q + fv0;
        F[1] = c->ad_c[1] * q + fv1;
    } else {
        F[0] = fv0;
        F[1] = fv1;
    }
}

/* ader_space_time_predictor, Advection-Diffusion/space_time_predictor.f90:10-155 */
void orc_predictor_ad(orc_ctx *c, double *lqhi, double *lFhi, double *lqbnd, double *lgradqbnd, const double *luh)
{
    const int n = c->n;
    const double *w = c->w;
    double lqh[NMAX * NMAX * NMAX], rhs0[NMAX * NMAX * NMAX], rhs[NMAX * NMAX * NMAX];
    double lqx[NMAX * NMAX * NMAX], lqy[NMAX * NMAX * NMAX], lFh[2 * NMAX * NMAX * NMAX];
    double lgx[NMAX * NMAX], lgy[NMAX * NMAX];
#define A3(i, j, l) ((i) + n * ((j) + n * (l)))
    for (int j = 0; j < n; j++) /* :42-58 */
        for (int i = 0; i < n; i++)
            for (int l = 0; l < n; l++) {
                lqh[A3(i, j, l)] = luh[i + n * j];
                rhs0[A3(i, j, l)] = w[i] * w[j] * c->f0[l] * luh[i + n * j];
            }
    OPS(3 * n * n * n);
    for (int iter = 0; iter < n; iter++) { /* :62-112 */
        for (int l = 0; l < n; l++) {
            OPS(2 * n * n * (2 * n + 2) + n * n * (c->ad_fixed ? 6 : 2) + 2 * n * (3 + n * (2 * n + 2)));
            for (int j = 0; j < n; j++) /* :67-69 */
                for (int i = 0; i < n; i++) {
                    double acc = 0.0;
                    for (int k = 0; k < n; k++) acc = acc + lqh[A3(k, j, l)] * c->dudx[i + n * k];
                    lqx[A3(i, j, l)] = 1.0 / c->dx[0] * acc;
                }
            for (int i = 0; i < n; i++) /* :73-75 */
                for (int j = 0; j < n; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < n; k++) acc = acc + lqh[A3(i, k, l)] * c->dudx[j + n * k];
                    lqy[A3(i, j, l)] = 1.0 / c->dx[1] * acc;
                }
            for (int j = 0; j < n; j++) /* :78-85 */
                for (int i = 0; i < n; i++) {
                    const double g[2] = {lqx[A3(i, j, l)], lqy[A3(i, j, l)]};
                    ad_flux(c, c->ad_fixed, lqh[A3(i, j, l)], g, &lFh[2 * A3(i, j, l)]);
                }
            for (int j = 0; j < n; j++) { /* :91-93 */
                const double sx = w[l] * w[j] * c->dt / c->dx[0];
                for (int i = 0; i < n; i++) {
                    double acc = 0.0;
                    for (int k = 0; k < n; k++) acc = acc + lFh[2 * A3(k, j, l)] * c->kxi[k + n * i];
                    rhs[A3(i, j, l)] = rhs0[A3(i, j, l)] - sx * acc;
                }
            }
            for (int i = 0; i < n; i++) { /* :97-99 */
                const double sy = w[l] * w[i] * c->dt / c->dx[1];
                for (int j = 0; j < n; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < n; k++) acc = acc + lFh[1 + 2 * A3(i, k, l)] * c->kxi[k + n * j];
                    rhs[A3(i, j, l)] = rhs[A3(i, j, l)] - sy * acc;
                }
            }
        }
        for (int j = 0; j < n; j++) /* :105-109 */
            for (int i = 0; i < n; i++) {
                const double f = 1.0 / (w[i] * w[j]);
                double tmp[NMAX];
                for (int l = 0; l < n; l++) {
                    double acc = 0.0;
                    for (int m = 0; m < n; m++) acc = acc + rhs[A3(i, j, m)] * c->ik1[l + n * m];
                    tmp[l] = f * acc;
                }
                for (int l = 0; l < n; l++) lqh[A3(i, j, l)] = tmp[l];
            }
        OPS(n * n * (2 + n * (2 * n + 1)));
    }
    OPS(5 * n * n * 2 * n + 2 * n * 6 * 2 * n);
    for (int j = 0; j < n; j++) /* :114-124 time averages */
        for (int i = 0; i < n; i++) {
            double a = 0.0, gx = 0.0, gy = 0.0, f0 = 0.0, f1 = 0.0;
            for (int l = 0; l < n; l++) a = a + lqh[A3(i, j, l)] * w[l];
            for (int l = 0; l < n; l++) gx = gx + lqx[A3(i, j, l)] * w[l];
            for (int l = 0; l < n; l++) gy = gy + lqy[A3(i, j, l)] * w[l];
            for (int l = 0; l < n; l++) f0 = f0 + lFh[2 * A3(i, j, l)] * w[l];
            for (int l = 0; l < n; l++) f1 = f1 + lFh[1 + 2 * A3(i, j, l)] * w[l];
            lqhi[i + n * j] = a;
            lgx[i + n * j] = gx;
            lgy[i + n * j] = gy;
            lFhi[0 + 2 * (i + n * j)] = f0;
            lFhi[1 + 2 * (i + n * j)] = f1;
        }
    for (int k = 0; k < 4 * n; k++) lqbnd[k] = 0.0;
    /* :130-152 boundary extrapolation; lqbnd(1,f,p), lgradqbnd(1,d,f,p) */
    for (int j = 0; j < n; j++) {
        double q0 = 0.0, q1 = 0.0, a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
        for (int i = 0; i < n; i++) q0 = q0 + lqhi[i + n * j] * c->flc[i];
        for (int i = 0; i < n; i++) q1 = q1 + lqhi[i + n * j] * c->frc[i];
        for (int i = 0; i < n; i++) a0 = a0 + lgx[i + n * j] * c->flc[i];
        for (int i = 0; i < n; i++) b0 = b0 + lgy[i + n * j] * c->flc[i];
        for (int i = 0; i < n; i++) a1 = a1 + lgx[i + n * j] * c->frc[i];
        for (int i = 0; i < n; i++) b1 = b1 + lgy[i + n * j] * c->frc[i];
        lqbnd[0 + 4 * j] = q0;
        lqbnd[1 + 4 * j] = q1;
        lgradqbnd[0 + 2 * (0 + 4 * j)] = a0;
        lgradqbnd[1 + 2 * (0 + 4 * j)] = b0;
        lgradqbnd[0 + 2 * (1 + 4 * j)] = a1;
        lgradqbnd[1 + 2 * (1 + 4 * j)] = b1;
    }
    for (int i = 0; i < n; i++) {
        double q0 = 0.0, q1 = 0.0, a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
        for (int j = 0; j < n; j++) q0 = q0 + lqhi[i + n * j] * c->flc[j];
        for (int j = 0; j < n; j++) q1 = q1 + lqhi[i + n * j] * c->frc[j];
        for (int j = 0; j < n; j++) a0 = a0 + lgx[i + n * j] * c->flc[j];
        for (int j = 0; j < n; j++) b0 = b0 + lgy[i + n * j] * c->flc[j];
        for (int j = 0; j < n; j++) a1 = a1 + lgx[i + n * j] * c->frc[j];
        for (int j = 0; j < n; j++) b1 = b1 + lgy[i + n * j] * c->frc[j];
        lqbnd[2 + 4 * i] = q0;
        lqbnd[3 + 4 * i] = q1;
        lgradqbnd[0 + 2 * (2 + 4 * i)] = a0;
        lgradqbnd[1 + 2 * (2 + 4 * i)] = b0;
        lgradqbnd[0 + 2 * (3 + 4 * i)] = a1;
        lgradqbnd[1 + 2 * (3 + 4 * i)] = b1;
    }
#undef A3
}

/* PDEViscLLFRiemannSolver, Advection-Diffusion/pde.f90:135-154 */
static double ad_llf(const orc_ctx *c, double QL, const double *gL, double QR, const double *gR, const double *nv)
{
    const double m_pi = 4.0 * atan(1.0);
    const double LL = c->ad_c[0] * nv[0] + c->ad_c[1] * nv[1], LR = LL;
    const double factor = (double)(2 * c->N + 1);
    const double eta = factor / (c->dx[0] * sqrt(0.5 * m_pi));
    const double smax_c = f_max(fabs(LL), fabs(LR));
    const double smax_v = f_max(c->ad_a, c->ad_a);
    const double smax = smax_c + 2.0 * eta * smax_v;
    double FL[2], FR[2];
    ad_flux(c, 1, QL, gL, FL);
    ad_flux(c, 1, QR, gR, FR);
    double mm = 0.0;
    mm = mm + (FL[0] + FR[0]) * nv[0];
    mm = mm + (FL[1] + FR[1]) * nv[1];
    return 0.5 * mm - 0.5 * smax * (QR - QL);
}

/* solve_riemann_problem, Advection-Diffusion/riemann_problem.f90:8-113 */
static void riemann_ad(orc_ctx *c)
{
    const int n = c->n, I = c->imax, J = c->jmax;
    const double m_pi = 4.0 * atan(1.0);
#define QB(ci, cj, f, p) (c->qbnd[(f) + 4 * ((p) + n * ((size_t)(ci) + (size_t)I * (cj)))])
#define GB(ci, cj, f, p) (&c->gradQbnd[2 * ((f) + 4 * ((p) + n * ((size_t)(ci) + (size_t)I * (cj))))])
    double nv[2] = {1.0, 0.0};
    for (int j = 0; j < J; j++)
        for (int i = 0; i <= I; i++)
            for (int l = 0; l < n; l++) {
                double QL, QR, gL[2], gR[2];
                const double *a, *b;
                if (i == 0) {
                    const int bc = c->ad_bc[0];
                    if (bc == 1) {
                        QL = QR = QB(0, j, 0, l);
                        a = b = GB(0, j, 0, l);
                    } else if (bc == 2) { /* :30-33: the state of the last cell, a reference quirk */
                        const double y = ((double)j + c->xi[l]) * c->dx[1];
                        QR = QB(I - 1, j, 1, l);
                        b = GB(I - 1, j, 1, l);
                        QL = -QR + 2.0 * sin(m_pi * y);
                        a = b;
                    } else {
                        QL = QB(I - 1, j, 1, l);
                        a = GB(I - 1, j, 1, l);
                        QR = QB(0, j, 0, l);
                        b = GB(0, j, 0, l);
                    }
                } else if (i == I) {
                    const int bc = c->ad_bc[1];
                    QL = QB(I - 1, j, 1, l);
                    a = GB(I - 1, j, 1, l);
                    if (bc == 1) {
                        QR = QL;
                        b = a;
                    } else if (bc == 2) {
                        QR = -QL;
                        b = a;
                    } else {
                        QR = QB(0, j, 0, l);
                        b = GB(0, j, 0, l);
                    }
                } else {
                    QL = QB(i - 1, j, 1, l);
                    a = GB(i - 1, j, 1, l);
                    QR = QB(i, j, 0, l);
                    b = GB(i, j, 0, l);
                }
                gL[0] = a[0]; gL[1] = a[1];
                gR[0] = b[0]; gR[1] = b[1];
                c->Fu[l + n * ((size_t)i + (size_t)(I + 1) * j)] = ad_llf(c, QL, gL, QR, gR, nv);
            }
    nv[0] = 0.0;
    nv[1] = 1.0;
    for (int j = 0; j <= J; j++)
        for (int i = 0; i < I; i++)
            for (int k = 0; k < n; k++) {
                double QL, QR, gL[2], gR[2];
                const double *a, *b;
                if (j == 0) {
                    const int bc = c->ad_bc[2];
                    if (bc == 1) {
                        QL = QR = QB(i, 0, 2, k);
                        a = b = GB(i, 0, 2, k);
                    } else if (bc == 2) {
                        QR = QB(i, 0, 2, k);
                        b = GB(i, 0, 2, k);
                        QL = -QR;
                        a = b;
                    } else {
                        QL = QB(i, J - 1, 3, k);
                        a = GB(i, J - 1, 3, k);
                        QR = QB(i, 0, 2, k);
                        b = GB(i, 0, 2, k);
                    }
                } else if (j == J) {
                    const int bc = c->ad_bc[3];
                    QL = QB(i, J - 1, 3, k);
                    a = GB(i, J - 1, 3, k);
                    if (bc == 1) {
                        QR = QL;
                        b = a;
                    } else if (bc == 2) {
                        QR = -QL;
                        b = a;
                    } else {
                        QR = QB(i, 0, 2, k);
                        b = GB(i, 0, 2, k);
                    }
                } else {
                    QL = QB(i, j - 1, 3, k);
                    a = GB(i, j - 1, 3, k);
                    QR = QB(i, j, 2, k);
                    b = GB(i, j, 2, k);
                }
                gL[0] = a[0]; gL[1] = a[1];
                gR[0] = b[0]; gR[1] = b[1];
                c->Gu[k + n * ((size_t)i + (size_t)I * j)] = ad_llf(c, QL, gL, QR, gR, nv);
            }
#undef QB
#undef GB
}

/* calc_time_step, Advection-Diffusion/calc_time_step.f90:9-57 */
static double calc_dt_ad(orc_ctx *c)
{
    const int n = c->n;
    const double factor = (double)(2 * c->N + 1);
    double dt = 1.0e20;
    for (size_t k = 0; k < ncell(c) * n * n; k++) {
        double denom = 0.0;
        for (int d = 0; d < 2; d++) {
            const double lam = c->ad_c[0] * (d == 0 ? 1.0 : 0.0) + c->ad_c[1] * (d == 0 ? 0.0 : 1.0);
            const double smax_c = fabs(lam), smax_v = fabs(c->ad_a);
            denom = denom + (smax_c + 2.0 * smax_v * factor / c->dx[d]) / c->dx[d];
        }
        dt = f_min(dt, c->cflpnpm / denom);
    }
    if (c->time + dt > c->tend) dt = c->tend - c->time;
    c->dt = dt;
    return dt;
}

/* one pass of Advection-Diffusion/main.f90:22-74 */
static int step_ad(orc_ctx *c)
{
    const int n = c->n, I = c->imax, J = c->jmax;
    const size_t blk = (size_t)n * n;
    calc_dt_ad(c);
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            const size_t k = (size_t)ci + (size_t)I * cj;
            orc_predictor_ad(c, &c->qhi[blk * k], &c->Fhi[2 * blk * k], &c->qbnd[4 * n * k], &c->gradQbnd[8 * n * k],
                             &c->uh[blk * k]);
        }
    for (size_t k = 0; k < ncell(c); k++) volume_impl(c, &c->duh[blk * k], &c->Fhi[2 * blk * k], NULL);
    riemann_ad(c);
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            const size_t k = (size_t)ci + (size_t)I * cj;
            orc_surface(c, &c->duh[blk * k], &c->Fu[n * (ci + (I + 1) * cj)], &c->Fu[n * (ci + 1 + (I + 1) * cj)],
                        &c->Gu[n * (ci + I * cj)], &c->Gu[n * (ci + I * (cj + 1))]);
        }
    for (size_t k = 0; k < ncell(c); k++) orc_update(c, &c->uh[blk * k], &c->duh[blk * k]);
    c->time = c->time + c->dt;
    return 0;
}

/* one pass of the time loop body, Euler/main.f90:28-100 (Baer-Nunziato/main.f90:28-100) */
int orc_step(orc_ctx *c)
{
    if (c->pde == ORC_PDE_AD) return step_ad(c);
    const int V = c->V, n = c->n, s = c->s, I = c->imax, J = c->jmax;
    size_t nc = ncell(c), blk = (size_t)V * n * n;
    int nrec = 0;
    orc_calc_dt(c);
    get_min_max(c);
    memset(c->recompute, 0, nc * sizeof(int));
    memcpy(c->olduh, c->uh, blk * nc * sizeof(double));
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            size_t k = (size_t)ci + (size_t)I * cj;
            predictor_impl(c, &c->qhi[blk * k], &c->Fhi[2 * blk * k], &c->Shi[blk * k], BND(c, c->qbnd, ci, cj),
                           BND(c, c->Fbnd, ci, cj), UH(c, c->uh, ci, cj));
        }
    const int bn = c->pde == ORC_PDE_BN;
    for (size_t k = 0; k < nc; k++) volume_impl(c, &c->duh[blk * k], &c->Fhi[2 * blk * k], bn ? &c->Shi[blk * k] : NULL);
    if (bn)
        riemann_bn(c);
    else
        riemann(c);
    /* BN: surface_integral(duh, Fm(i), Fp(i+1), Gm(j), Gp(j+1)), Baer-Nunziato/main.f90:66-70 */
    const double *FR_ = bn ? c->Fp : c->Fu, *GT_ = bn ? c->Gp : c->Gu;
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            size_t k = (size_t)ci + (size_t)I * cj;
            orc_surface(c, &c->duh[blk * k], &c->Fu[V * n * (ci + (I + 1) * cj)], &FR_[V * n * (ci + 1 + (I + 1) * cj)],
                        &c->Gu[V * n * (ci + I * cj)], &GT_[V * n * (ci + I * (cj + 1))]);
        }
    for (size_t k = 0; k < nc; k++) {
        orc_update(c, &c->uh[blk * k], &c->duh[blk * k]);
        if (c->margin) c->margin[k] = orc_dmp_margin(c, &c->uh[blk * k], &c->lmin[V * k], &c->lmax[V * k], 0.0);
        if (!orc_dmp(c, &c->uh[blk * k], &c->lmin[V * k], &c->lmax[V * k], 0.0)) {
            c->recompute[k] = 1;
            nrec++;
        }
    }
    /* spread_recompute: Euler/limiter.f90:329-354 */
    static const int f2n[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int cj = 0; cj < J; cj++)
        for (int ci = 0; ci < I; ci++) {
            if (c->recompute[(size_t)ci + (size_t)I * cj] != 1) continue;
            for (int f = 0; f < 4; f++) {
                int in, jn;
                neighbor(c, ci, cj, f2n[f][0], f2n[f][1], &in, &jn);
                size_t kn = (size_t)in + (size_t)I * jn;
                if (c->recompute[kn] == 0) c->recompute[kn] = 2;
            }
        }
    /* allocate_limiter: Euler/limiter.f90:302-322 */
    for (size_t k = 0; k < nc; k++) c->status[k] = c->recompute[k] == 1 ? 1 : 0;
    if (bn)
        subcell_recompute_bn(c);
    else
        subcell_recompute(c);
    /* update_limiter: Euler/limiter.f90:272-296 */
    for (size_t k = 0; k < nc; k++) {
        if (c->status[k] != 0) memcpy(&c->Lh[(size_t)V * s * s * k], &c->NewLh[(size_t)V * s * s * k], sizeof(double) * V * s * s);
        c->oldstatus[k] = c->status[k];
    }
    c->time = c->time + c->dt;
    return nrec;
}
