// operators.hpp -- ADER-DG operator matrices (host), see operators.cpp.
#pragma once

namespace aderdg {

constexpr int kMaxDegree = 4;               // highest N instantiated on the device
constexpr int kMaxN = kMaxDegree + 1;       // nodes per direction
constexpr int kMaxS = 2 * kMaxDegree + 1;   // subcells per direction

// Column-major, the layouts of Euler/ader_dg.f90:30-38,63-64 with n = N+1, s = 2N+1.
struct Operators {
    int N;
    double xi[kMaxN], w[kMaxN];
    double kxi[kMaxN * kMaxN];   // Kxi(k,l) = w_l phi'_k(xi_l)
    double dudx[kMaxN * kMaxN];  // iMM * Kxi^T
    double ik1[kMaxN * kMaxN];   // (phi(1) phi(1)^T - Kxi)^-1
    double f0[kMaxN], flc[kMaxN], frc[kMaxN];
    double u2l[kMaxS * kMaxN];   // uh2lim(s,n)
    double l2u[kMaxN * kMaxS];   // lim2uh(n,s)
    double u2b[kMaxN * kMaxN];   // uh2lob(n,n)
};

// Returns false for an unsupported degree or a singular matrix.
bool build_operators(int N, Operators &op);

}  // namespace aderdg
