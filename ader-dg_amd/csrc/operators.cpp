// operators.cpp -- host construction of the ADER-DG operator matrices.
//
// Restates the setup of Euler/initialize.f90:80-213 (with gauleg/gaulob of
// Euler/quadrature.f90:8-101, basis_func_1d of Euler/basis.f90:9-41 and
// MatrixInverse of Euler/la.f90:8-65) so that callers without a Fortran host
// (the batched engine, the serving loop, tests) get the same matrices the
// reference module holds.  Compiled without floating-point contraction; the
// CPU tests check every matrix bit-for-bit against the reference's.
#include "operators.hpp"

#pragma clang fp contract(off)

#include <cmath>
#include <cstring>

namespace aderdg {

namespace {

void gauss_legendre(double x1, double x2, double *x, double *w, int n)
{
    const double eps = 1.0e-15, pi = 4.0 * std::atan(1.0);
    const int m = (n + 1) / 2;
    const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = std::cos(pi * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
        do {  // Newton on P_n
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / static_cast<double>(j);
            }
            pp = static_cast<double>(n) * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > eps);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

void gauss_lobatto(double x1, double x2, double *x, double *w, int n1)
{
    const double eps = 1.0e-15, pi = 4.0 * std::atan(1.0);
    const int nd = n1 - 1;
    double xold[kMaxN], P[kMaxN][kMaxN + 1];
    for (int i = 0; i <= nd; i++) x[i] = std::cos(pi * static_cast<double>(i) / static_cast<double>(nd));
    for (int iter = 0; iter < 100000; iter++) {
        for (int a = 0; a < n1; a++) xold[a] = x[a];
        for (int a = 0; a < n1; a++) {
            P[a][1] = 1.0;
            P[a][2] = x[a];
            for (int k = 2; k <= nd; k++)
                P[a][k + 1] = (static_cast<double>(2 * k - 1) * x[a] * P[a][k] - static_cast<double>(k - 1) * P[a][k - 1]) /
                              static_cast<double>(k);
        }
        double dmax = 0.0;
        for (int a = 0; a < n1; a++) {
            x[a] = xold[a] - (x[a] * P[a][n1] - P[a][nd]) / (static_cast<double>(n1) * P[a][n1]);
            dmax = std::fmax(dmax, std::fabs(x[a] - xold[a]));
        }
        if (dmax <= eps) break;
    }
    for (int a = 0; a < n1; a++) {
        w[a] = 2.0 / (static_cast<double>(nd * n1) * (P[a][n1] * P[a][n1]));
        w[a] = 0.5 * w[a] * (x2 - x1);
        x[a] = x1 + 0.5 * (x[a] + 1.0) * (x2 - x1);
    }
}

// Lagrange basis through the nodes xin and its derivative at xi
void lagrange(const double *xin, int n, double xi, double *phi, double *dphi)
{
    for (int m = 0; m < n; m++) {
        phi[m] = 1.0;
        dphi[m] = 0.0;
    }
    for (int m = 0; m < n; m++) {
        for (int j = 0; j < n; j++)
            if (j != m) phi[m] = phi[m] * (xi - xin[j]) / (xin[m] - xin[j]);
        for (int i = 0; i < n; i++) {
            if (i == m) continue;
            double t = 1.0;
            for (int j = 0; j < n; j++)
                if (j != i && j != m) t = t * (xi - xin[j]) / (xin[m] - xin[j]);
            dphi[m] = dphi[m] + t / (xin[m] - xin[i]);
        }
    }
}

// Gauss-Jordan inverse working on the transpose with column pivoting, the
// elimination order of Euler/la.f90:8-65.  Column-major NxN.
bool invert(int n, const double *A, double *iA)
{
    double C[2 * (kMaxN + 1)][kMaxN + 1];
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) C[r][c] = A[c + n * r];
    for (int r = n; r < 2 * n; r++)
        for (int c = 0; c < n; c++) C[r][c] = (r - n == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; i++) {
        int piv = i;
        for (int c = i + 1; c < n; c++)
            if (std::fabs(C[i][c]) > std::fabs(C[i][piv])) piv = c;
        for (int r = 0; r < 2 * n; r++) std::swap(C[r][piv], C[r][i]);
        if (std::fabs(C[i][i]) <= 1.0e-15) return false;
        const double inv = 1.0 / C[i][i];
        for (int r = 0; r < 2 * n; r++) C[r][i] = C[r][i] * inv;
        for (int c = i + 1; c < n; c++) {
            const double f = C[i][c];
            for (int r = 0; r < 2 * n; r++) C[r][c] = C[r][c] - f * C[r][i];
        }
    }
    for (int i = n - 1; i >= 0; i--)
        for (int c = i - 1; c >= 0; c--) {
            const double f = C[i][c];
            for (int r = 0; r < 2 * n; r++) C[r][c] = C[r][c] - f * C[r][i];
        }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) iA[r + n * c] = C[n + c][r];
    return true;
}

}  // namespace

bool build_operators(int N, Operators &op)
{
    const int n = N + 1, s = 2 * N + 1;
    if (N < 1 || n > kMaxN) return false;
    std::memset(&op, 0, sizeof op);
    op.N = N;
    double phi[kMaxN], dphi[kMaxN], phi0[kMaxN], phi1[kMaxN], mm[kMaxN * kMaxN], k1[kMaxN * kMaxN];
    gauss_legendre(0.0, 1.0, op.xi, op.w, n);
    std::memset(mm, 0, sizeof mm);
    for (int g = 0; g < n; g++) {
        lagrange(op.xi, n, op.xi[g], phi, dphi);
        for (int k = 0; k < n; k++)
            for (int l = 0; l < n; l++) {
                mm[k + n * l] = mm[k + n * l] + op.w[g] * phi[k] * phi[l];
                op.kxi[k + n * l] = op.kxi[k + n * l] + op.w[g] * dphi[k] * phi[l];
            }
    }
    double imm[kMaxN * kMaxN];
    if (!invert(n, mm, imm)) return false;
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc = acc + imm[a + n * k] * op.kxi[b + n * k];
            op.dudx[a + n * b] = acc;
        }
    lagrange(op.xi, n, 0.0, phi0, dphi);
    lagrange(op.xi, n, 1.0, phi1, dphi);
    for (int k = 0; k < n; k++) {
        op.f0[k] = phi0[k];
        op.flc[k] = phi0[k];
        op.frc[k] = phi1[k];
        for (int l = 0; l < n; l++) k1[k + n * l] = phi1[k] * phi1[l] - op.kxi[k + n * l];
    }
    if (!invert(n, k1, op.ik1)) return false;
    // subcell averages and the average-preserving least-squares way back
    const double dxi = 1.0 / static_cast<double>(s);
    for (int i = 1; i <= s; i++) {
        const double xi1 = static_cast<double>(i - 1) * dxi;
        for (int ii = 0; ii < n; ii++) {
            lagrange(op.xi, n, xi1 + dxi * op.xi[ii], phi, dphi);
            for (int a = 0; a < n; a++) op.u2l[(i - 1) + s * a] = op.u2l[(i - 1) + s * a] + op.w[ii] * phi[a];
        }
    }
    const int m = n + 1;
    double lsq[(kMaxN + 1) * (kMaxN + 1)], ilsq[(kMaxN + 1) * (kMaxN + 1)], rhs[(kMaxN + 1) * kMaxS];
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            double acc = 0.0;
            for (int p = 0; p < s; p++) acc = acc + op.u2l[p + s * a] * op.u2l[p + s * b];
            lsq[a + m * b] = 2.0 * acc;
        }
    for (int a = 0; a < n; a++) {
        lsq[n + m * a] = op.w[a];
        lsq[a + m * n] = -op.w[a];
    }
    lsq[n + m * n] = 0.0;
    if (!invert(m, lsq, ilsq)) return false;
    for (int a = 0; a < n; a++)
        for (int p = 0; p < s; p++) rhs[a + m * p] = 2.0 * op.u2l[p + s * a];
    for (int p = 0; p < s; p++) rhs[n + m * p] = dxi;
    for (int a = 0; a < n; a++)
        for (int p = 0; p < s; p++) {
            double acc = 0.0;
            for (int k = 0; k < m; k++) acc = acc + ilsq[a + m * k] * rhs[k + m * p];
            op.l2u[a + n * p] = acc;
        }
    double xlob[kMaxN], wlob[kMaxN];
    gauss_lobatto(0.0, 1.0, xlob, wlob, n);
    for (int ii = 0; ii < n; ii++) {
        lagrange(op.xi, n, xlob[ii], phi, dphi);
        for (int a = 0; a < n; a++) op.u2b[ii + n * a] = phi[a];
    }
    return true;
}

}  // namespace aderdg
