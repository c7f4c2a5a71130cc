// ad_kernels.hip -- CDNA4 (gfx950) kernels of the ADER-DG step for 2-D linear
// advection-diffusion (nVar = 1, Advection-Diffusion/), over solve groups like
// the Euler / Baer-Nunziato kernels of step_kernels.hip:
//   k_predictor_ad  ader_space_time_predictor (with the reference's viscous-only
//                   predictor flux or the fused one) + volume_integral + face
//                   extrapolation of q and grad q
//   k_face_ad       solve_riemann_problem: viscous LLF with Neumann, Dirichlet
//                   and periodic boundaries, one thread per face point
//   k_update_ad     surface_integral + element_update + the next step's
//                   time-step candidate (the system has no limiter)
// k_dt, k_solves and k_init of step_kernels.hip complete the step.
//
// Face data (cell arrays strided by G): fq [face][p][cell] = q,
// ff [face][d][p][cell] = dq/dx_d; xflux/yflux [p][face].
#include "pde_device.hpp"
#include "step_kernels.hpp"

namespace aderdg {

namespace {

constexpr int kAdBlock = 256;

template <int N>
struct AdShape {
    static constexpr int n = N + 1, NN = n * n;
    static constexpr int CPW = kAdBlock / NN;  // cells per workgroup
};

// PDEViscFlux (Advection-Diffusion/pde.f90:54-64) or PDEFusedViscConvFlux (:115-129)
ADG_DEV void ad_flux(const MeshArgs &m, bool fused, double q, double gx, double gy, double *fx, double *fy)
{
    const double vx = -(m.diff * gx), vy = -(m.diff * gy);
    if (fused) {
        *fx = m.adv[0] * q + vx;
        *fy = m.adv[1] * q + vy;
    } else {
        *fx = vx;
        *fy = vy;
    }
}

// ---------------------------------------------------------------------------
// Thread (cell, i, j) owns the time line of its node.  Per time slice: q goes to
// LDS, the thread forms its gradient (dudx contractions), evaluates the flux and
// writes it to LDS, then contracts its row of Fx and column of Fy with Kxi (two
// barriers per slice); the time contraction with iK1 is thread-local.
// Advection-Diffusion/space_time_predictor.f90:10-155.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(kAdBlock) void k_predictor_ad(StepBufs b)
{
    using S = AdShape<N>;
    constexpr int n = S::n, NN = S::NN, CPW = S::CPW;
    __shared__ double sq[CPW][NN], sfx[CPW][NN], sfy[CPW][NN];
    __shared__ double sgx[CPW][NN], sgy[CPW][NN];
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const bool own = lc < CPW;
    const int lcs = own ? lc : CPW - 1;
    if (blockIdx.x == 0 && tid == 0) {
        b.counters[kCntList] = 0;
        b.counters[kCntTroubled] = 0;
    }
    const double wi = op->w[i], wj = op->w[j];
    double kx_i[n], kx_j[n], du_i[n], du_j[n];
#pragma unroll
    for (int k = 0; k < n; k++) {
        kx_i[k] = op->kxi[k + n * i];
        kx_j[k] = op->kxi[k + n * j];
        du_i[k] = op->dudx[i + n * k];
        du_j[k] = op->dudx[j + n * k];
    }
    const int ntab = ld_const(b.meta + kTabPred);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabPred] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const SolveArgs sa = ld_const(b.solves + w.x);
        const MeshArgs &m = sa.m;
        const bool fused = m.ad_fixed != 0;
        const int cell = w.y + lc;
        const size_t g = (size_t)sa.cell0 + cell;
        const bool act = own && cell < m.ncell;
        const double dt = st.dt;
        const double u = act ? b.uh[st.parity][g * NN + tt] : 0.0;
        double q[n];
#pragma unroll
        for (int l = 0; l < n; l++) q[l] = u;
        double qh = 0.0, fxh = 0.0, fyh = 0.0, gxh = 0.0, gyh = 0.0;
#pragma unroll 1
        for (int it = 0; it <= N; it++) {  // Picard iterations, :62-112
            double qn[n];
#pragma unroll
            for (int ll = 0; ll < n; ll++) qn[ll] = 0.0;
#pragma unroll
            for (int l = 0; l < n; l++) {
                if (own) sq[lcs][tt] = q[l];
                __syncthreads();
                double ax = 0.0, ay = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) ax = ax + sq[lcs][k + n * j] * du_i[k];
#pragma unroll
                for (int k = 0; k < n; k++) ay = ay + sq[lcs][i + n * k] * du_j[k];
                const double gx = 1.0 / m.dx * ax, gy = 1.0 / m.dy * ay;  // :67-75
                double fx, fy;
                ad_flux(m, fused, q[l], gx, gy, &fx, &fy);  // :78
                if (own) {
                    sfx[lcs][j * n + i] = fx;
                    sfy[lcs][i * n + j] = fy;
                }
                const double wl = op->w[l];
                if (it == N) {  // time averages of the last sweep, :114-124
                    fxh = fxh + fx * wl;
                    fyh = fyh + fy * wl;
                    gxh = gxh + gx * wl;
                    gyh = gyh + gy * wl;
                }
                __syncthreads();
                const double sx = wl * wj * dt / m.dx, sy = wl * wi * dt / m.dy;
                const double c0 = wi * wj * op->f0[l];
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sfx[lcs][j * n + k] * kx_i[k];
                const double rr = c0 * u - sx * acc;  // :91-93
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sfy[lcs][i * n + k] * kx_j[k];
                const double rv = rr - sy * acc;  // :97-99
#pragma unroll
                for (int ll = 0; ll < n; ll++) qn[ll] = qn[ll] + rv * op->ik1[ll + n * l];
            }
            const double f = 1.0 / (wi * wj);  // :105-109
#pragma unroll
            for (int ll = 0; ll < n; ll++) q[ll] = f * qn[ll];
        }
#pragma unroll
        for (int l = 0; l < n; l++) qh = qh + q[l] * op->w[l];
        // volume integral (Advection-Diffusion/volume_integral.f90:8-36) and face extrapolation
        __syncthreads();
        if (own) {
            sq[lcs][tt] = qh;
            sfx[lcs][tt] = fxh;
            sfy[lcs][tt] = fyh;
            sgx[lcs][tt] = gxh;
            sgy[lcs][tt] = gyh;
        }
        __syncthreads();
        {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) acc = acc + sfx[lcs][k + n * j] * op->kxi[i + n * k];
            double d = 0.0 + acc * wj / m.dx;
            acc = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) acc = acc + sfy[lcs][i + n * k] * op->kxi[j + n * k];
            d = d + acc * wi / m.dy;
            if (act) b.lduh[g * NN + tt] = d;
        }
        // :130-152, one (face, point) per thread
        for (int task = tt; task < 4 * n; task += NN) {
            const int f = task / n, p = task - f * n;
            const OpPtr o2 = op;
            double aq = 0.0, agx = 0.0, agy = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) {
                const int idx = (f < 2) ? (k + n * p) : (p + n * k);
                const double cf = (f & 1) ? o2->frc[k] : o2->flc[k];
                aq = aq + sq[lcs][idx] * cf;
            }
#pragma unroll
            for (int k = 0; k < n; k++) {
                const int idx = (f < 2) ? (k + n * p) : (p + n * k);
                const double cf = (f & 1) ? o2->frc[k] : o2->flc[k];
                agx = agx + sgx[lcs][idx] * cf;
            }
#pragma unroll
            for (int k = 0; k < n; k++) {
                const int idx = (f < 2) ? (k + n * p) : (p + n * k);
                const double cf = (f & 1) ? o2->frc[k] : o2->flc[k];
                agy = agy + sgy[lcs][idx] * cf;
            }
            if (act) {
                b.fq[((size_t)f * n + p) * b.G + g] = aq;
                b.ff[((size_t)(f * 2 + 0) * n + p) * b.G + g] = agx;
                b.ff[((size_t)(f * 2 + 1) * n + p) * b.G + g] = agy;
            }
        }
        __syncthreads();
    }
}

// PDEViscLLFRiemannSolver, Advection-Diffusion/pde.f90:135-154
ADG_DEV double ad_llf(const MeshArgs &m, int N, double QL, const double *gL, double QR, const double *gR, double nx,
                      double ny)
{
    const double m_pi = 3.141592653589793;
    const double LL = m.adv[0] * nx + m.adv[1] * ny, LR = LL;
    const double factor = (double)(2 * N + 1);
    const double eta = factor / (m.dx * sqrt(0.5 * m_pi));
    const double smax_c = f_max(fabs(LL), fabs(LR));
    const double smax_v = f_max(m.diff, m.diff);
    const double smax = smax_c + 2.0 * eta * smax_v;
    double FL0, FL1, FR0, FR1;
    ad_flux(m, true, QL, gL[0], gL[1], &FL0, &FL1);
    ad_flux(m, true, QR, gR[0], gR[1], &FR0, &FR1);
    double mm = 0.0;
    mm = mm + (FL0 + FR0) * nx;
    mm = mm + (FL1 + FR1) * ny;
    return 0.5 * mm - 0.5 * smax * (QR - QL);
}

// solve_riemann_problem, Advection-Diffusion/riemann_problem.f90:8-113: one thread per face point
template <int N>
__global__ __launch_bounds__(kBlock) void k_face_ad(StepBufs b)
{
    constexpr int n = N + 1;
    const int ntab = ld_const(b.meta + kTabFace);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabFace] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const SolveArgs sa = ld_const(b.solves + w.x);
        const MeshArgs &m = sa.m;
        const int I = m.imax, J = m.jmax;
        const long nfx = (long)(I + 1) * J, nfy = (long)I * (J + 1);
        const long t = (long)w.y + threadIdx.x;
        if (t >= (nfx + nfy) * n) continue;
        const bool xdir = t < nfx * n;
        const long t2 = xdir ? t : t - nfx * n;
        const long nf = xdir ? nfx : nfy;
        const int p = (int)(t2 / nf);
        const long fidx = t2 - (long)p * nf;
        const size_t c0 = sa.cell0;
        // q and grad q of face f of local cell c at point p
        auto ld = [&](int f, long c, double *Q, double *G) {
            *Q = b.fq[((size_t)f * n + p) * b.G + c0 + c];
            G[0] = b.ff[((size_t)(f * 2 + 0) * n + p) * b.G + c0 + c];
            G[1] = b.ff[((size_t)(f * 2 + 1) * n + p) * b.G + c0 + c];
        };
        double QL, QR, gL[2], gR[2], out;
        if (xdir) {
            const int fi = (int)(fidx % (I + 1)), j = (int)(fidx / (I + 1));
            if (fi == 0) {
                const int bc = m.ad_bc[0];
                if (bc == 1) {
                    ld(0, (long)I * j, &QL, gL);
                    QR = QL;
                    gR[0] = gL[0];
                    gR[1] = gL[1];
                } else if (bc == 2) {  // :30-33, with the last cell's state as the reference does
                    const double y = ((double)j + ((OpPtr)b.op)->xi[p]) * m.dy;
                    ld(1, (I - 1) + (long)I * j, &QR, gR);
                    QL = -QR + 2.0 * sin(3.141592653589793 * y);
                    gL[0] = gR[0];
                    gL[1] = gR[1];
                } else {
                    ld(1, (I - 1) + (long)I * j, &QL, gL);
                    ld(0, (long)I * j, &QR, gR);
                }
            } else if (fi == I) {
                const int bc = m.ad_bc[1];
                ld(1, (I - 1) + (long)I * j, &QL, gL);
                if (bc == 3) {
                    ld(0, (long)I * j, &QR, gR);
                } else {
                    QR = bc == 2 ? -QL : QL;
                    gR[0] = gL[0];
                    gR[1] = gL[1];
                }
            } else {
                ld(1, (fi - 1) + (long)I * j, &QL, gL);
                ld(0, fi + (long)I * j, &QR, gR);
            }
            out = ad_llf(m, N, QL, gL, QR, gR, 1.0, 0.0);
        } else {
            const int i = (int)(fidx % I), fj = (int)(fidx / I);
            if (fj == 0) {
                const int bc = m.ad_bc[2];
                ld(2, i, &QR, gR);
                if (bc == 3) {
                    ld(3, i + (long)I * (J - 1), &QL, gL);
                } else {
                    QL = bc == 2 ? -QR : QR;
                    gL[0] = gR[0];
                    gL[1] = gR[1];
                }
            } else if (fj == J) {
                const int bc = m.ad_bc[3];
                ld(3, i + (long)I * (J - 1), &QL, gL);
                if (bc == 3) {
                    ld(2, i, &QR, gR);
                } else {
                    QR = bc == 2 ? -QL : QL;
                    gR[0] = gL[0];
                    gR[1] = gL[1];
                }
            } else {
                ld(3, i + (long)I * (fj - 1), &QL, gL);
                ld(2, i + (long)I * fj, &QR, gR);
            }
            out = ad_llf(m, N, QL, gL, QR, gR, 0.0, 1.0);
        }
        double *dst = xdir ? b.xflux : b.yflux;
        const size_t stride = xdir ? b.GX : b.GY;
        dst[(size_t)p * stride + (size_t)(xdir ? sa.fx0 : sa.fy0) + fidx] = out;
    }
}

// time-step candidate of a node, Advection-Diffusion/calc_time_step.f90:36-48 (the same at every node)
ADG_DEV double ad_node_dt(const MeshArgs &m, int N)
{
    const double factor = (double)(2 * N + 1);
    double denom = 0.0;
    for (int d = 0; d < 2; d++) {
        const double lam = m.adv[0] * (d == 0 ? 1.0 : 0.0) + m.adv[1] * (d == 0 ? 0.0 : 1.0);
        const double h = d == 0 ? m.dx : m.dy;
        denom = denom + (fabs(lam) + 2.0 * fabs(m.diff) * factor / h) / h;
    }
    return m.cflpnpm / denom;
}

// surface_integral + element_update (Advection-Diffusion/surface_integral.f90:8-39,
// element_update.f90:8-30) and the next step's dt candidate; SEED: the candidate only.
template <int N, bool SEED>
__global__ __launch_bounds__(kAdBlock) void k_update_ad(StepBufs b)
{
    using S = AdShape<N>;
    constexpr int n = S::n, NN = S::NN, CPW = S::CPW;
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const int ntab = ld_const(b.meta + kTabUpd);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabUpd] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (SEED ? !st.seed : !st.active) continue;
        const SolveArgs sa = ld_const(b.solves + w.x);
        const MeshArgs &m = sa.m;
        const int cell = w.y + lc;
        const size_t g = (size_t)sa.cell0 + cell;
        if (lc >= CPW || cell >= m.ncell) continue;
        if (!SEED) {
            const int ci = cell % m.imax, cj = cell / m.imax;
            const size_t fl = sa.fx0 + ci + (size_t)(m.imax + 1) * cj, fb = sa.fy0 + ci + (size_t)m.imax * cj;
            const double wi = op->w[i], wj = op->w[j];
            double d = b.lduh[g * NN + tt];
            const double Fl = b.xflux[(size_t)j * b.GX + fl], Fr = b.xflux[(size_t)j * b.GX + fl + 1];
            const double Gb = b.yflux[(size_t)i * b.GY + fb], Gt = b.yflux[(size_t)i * b.GY + fb + m.imax];
            d = d - wj / m.dx * (Fr * op->frc[i] - Fl * op->flc[i]);
            d = d - wi / m.dy * (Gt * op->frc[j] - Gb * op->flc[j]);
            d = d / (wi * wj);
            b.uh[st.parity ^ 1][g * NN + tt] = b.uh[st.parity][g * NN + tt] + st.dt * d;
            if (tt == 0) b.rc[g] = 0;
        }
        if (tt == 0) b.cell_dt[g] = ad_node_dt(m, N);
    }
}

template <int N>
hipError_t launch_ad_impl(const StepBufs &b, hipStream_t st, int part, const int grid[kNumTabs])
{
    switch (part) {
    case kPartPredictor:
        hipLaunchKernelGGL((k_predictor_ad<N>), dim3(grid[kTabPred]), dim3(kAdBlock), 0, st, b);
        break;
    case kPartFace:
        hipLaunchKernelGGL((k_face_ad<N>), dim3(grid[kTabFace]), dim3(kBlock), 0, st, b);
        break;
    case kPartUpdate:
        hipLaunchKernelGGL((k_update_ad<N, false>), dim3(grid[kTabUpd]), dim3(kAdBlock), 0, st, b);
        break;
    case kPartSeed:
        hipLaunchKernelGGL((k_update_ad<N, true>), dim3(grid[kTabUpd]), dim3(kAdBlock), 0, st, b);
        break;
    case kPartSpread:
    case kPartSubcell:
    case kPartRanges:
        return hipSuccess;  // no limiter in the advection-diffusion solver
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

void table_quanta_ad(int N, int q[kNumTabs])
{
    const int NN = (N + 1) * (N + 1);
    q[kTabPred] = kAdBlock / NN;
    q[kTabUpd] = kAdBlock / NN;
    q[kTabFace] = kBlock;
    q[kTabCell] = kBlock;
}

hipError_t launch_part_ad(int N, const StepBufs &b, hipStream_t st, int part, const int grid[kNumTabs])
{
    switch (N) {
    case 1: return launch_ad_impl<1>(b, st, part, grid);
    case 2: return launch_ad_impl<2>(b, st, part, grid);
    case 3: return launch_ad_impl<3>(b, st, part, grid);
    case 4: return launch_ad_impl<4>(b, st, part, grid);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace aderdg
