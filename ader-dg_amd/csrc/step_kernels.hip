// step_kernels.hip -- CDNA4 (gfx950) kernels of the ADER-DG element step for the
// 2-D Euler (nVar = 4) and Baer-Nunziato (nVar = 9) systems.
//
// Per time step (the order of Euler/main.f90:24-100, Baer-Nunziato/main.f90:24-100):
//   k_dt          calc_time_step: per-solve min-reduction of per-cell candidates
//   k_ranges      the 3x3 pass of get_min_max (the DMP bounds of the step)
//   k_predictor   ader_space_time_predictor + volume_integral + face
//                 extrapolation
//   k_face        solve_riemann_problem: one thread per face point
//   k_update      surface_integral + element_update + DMP, and the next
//                 step's per-cell get_min_max pass and time-step candidates
//   k_spread      spread_recompute + allocate_limiter + wave-ballot
//                 compaction of the cells to recompute
//   k_subcell     subcell_recompute (MUSCL-Hancock on the (2N+1)^2 subgrid)
//                 + put_subcell_data + re-DMP + update_limiter, over the
//                 compacted list only (k_subcell_bn for Baer-Nunziato)
//
// All kernels serve a group of independent solves of one (PDE, N): a solve's
// cells are a contiguous range of the flat cell arrays (local cell
// i + IMAX*j at cell0 + i + IMAX*j) and its faces ranges of the face arrays.
// Data layout in HBM (G = cell capacity of the group):
//   uh, lduh      [cell][j][i][v]   the reference's uh(V,n,n,IMAX,JMAX)
//   fq, ff        [face][v][p][cell] face extrapolations of q and F*n,
//                 structure of arrays so that the face pass reads
//                 consecutive cells with consecutive lanes
//   xflux, yflux  [side][v][p][face]
//   lmin.. etc.   [v][cell]
//   Lh            [cell][q][p][v]   subcell averages of limited cells
// The step kernels loop over work tables of (solve, first item) entries
// (persistent workgroups); solves that are not active are skipped, and each
// solve has its own ping-pong parity, time and dt.
#include <type_traits>

#include "step_kernels.hpp"
#include "pde_device.hpp"

namespace aderdg {

namespace {

// workgroup sizes: Baer-Nunziato keeps three times the data per node
template <class P>
constexpr int kPredBlock = P::V == 4 ? 256 : 64;
// the matrix-core predictor sums in the matrix unit's order: product build only
#if defined(ADERDG_EXACT)
constexpr bool kMfmaBuild = false;
#else
constexpr bool kMfmaBuild = true;
#endif
template <class P>
constexpr int kUpdBlock = P::V == 4 ? 128 : 64;
// threads of the Baer-Nunziato subcell workgroup (one listed cell each): one interior face per thread
constexpr int kSubBnBlock = 128;
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// copies of the update kernel's LDS range slots per cell (k_update): 4 measured best of
// 1, 2, 4, 8, 16 (fewer lanes meet on one address; more copies cost occupancy)
constexpr int kUpdKeyCopies = 4;

template <class P, int N>
struct Shape {
    static constexpr int V = P::V;
    static constexpr int n = N + 1;
    static constexpr int NN = n * n;
    static constexpr int s = 2 * N + 1;
    static constexpr int SS = s * s;
    static constexpr int PRED_BLOCK = kPredBlock<P>;
    static constexpr int UPD_BLOCK = kUpdBlock<P>;
    static constexpr int CPW = PRED_BLOCK / NN;   // cells per predictor workgroup
    static constexpr int CPWU = UPD_BLOCK / NN;   // cells per update workgroup
    // predictor LDS per cell and time slice: Fx, Fy (and q for the NCP gradients)
    static constexpr int SLOT = (P::kNcp ? 3 : 2) * V * NN;
    // range slots of the update per cell: (min, max) of nodal / Lobatto / subcell values per v, dt
    static constexpr int NPART = 6 * V + 1;
};

template <class M>
ADG_DEV double dmr_shock_x(const M &m, double time) { return m.dmr_a + m.dmr_b * time; }



// ---------------------------------------------------------------------------
// space-time predictor (+ volume integral, face extrapolation, min/max pass 2)
// Thread (cell, i, j) owns the time line q(:, i, j, 0..N) of one spatial node.
// The x- and y-contractions read the fluxes of the other nodes of the same
// time slice from LDS (double buffered, one barrier per slice); the time
// contraction with iK1 is thread-local.  Baer-Nunziato also exchanges q for
// the gradients of the non-conservative product.
// ---------------------------------------------------------------------------
// Volume integral and boundary extrapolation of one
// cell from its time averages in LDS (qhiL, fhxL, fhyL, shL: [v][j][i]; shL is
// read for Baer-Nunziato only), spread over the cell's nthr threads.  TEST:
// write the averages and boundary values in the reference layouts instead.
template <class P, int N, bool TEST, class M>
ADG_DEV void predictor_finish(const M &m, const StepBufs &b, const TestBufs &t, const double *qhiL,
                              const double *fhxL, const double *fhyL, const double *shL, int tt, int nthr,
                              bool act, int cell, int cell0)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, V = S::V;
    const OpPtr op = (OpPtr)b.op;
    const size_t g = (size_t)cell0 + cell;
    for (int node = tt; act && node < NN; node += nthr) {
        if (TEST) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                t.qhi[(size_t)cell * V * NN + v + V * node] = qhiL[v * NN + node];
                t.fhi[(size_t)cell * 2 * V * NN + v + V * (0 + 2 * node)] = fhxL[v * NN + node];
                t.fhi[(size_t)cell * 2 * V * NN + v + V * (1 + 2 * node)] = fhyL[v * NN + node];
                if (P::kNcp) t.shi[(size_t)cell * V * NN + v + V * node] = shL[v * NN + node];
            }
        } else {
            // volume integral, Euler/volume_integral.f90:17-28 (+ Baer-Nunziato/volume_integral.f90:36-40)
            const int i = node % n, j = node / n;
            const double wi = op->w[i], wj = op->w[j];
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + fhxL[v * NN + k + n * j] * op->kxi[i + n * k];
                double d = 0.0 + acc * wj / m.dx;
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + fhyL[v * NN + i + n * k] * op->kxi[j + n * k];
                d = d + acc * wi / m.dy;
                if (P::kNcp) d = d + wi * wj * shL[v * NN + node];
                b.lduh[g * V * NN + V * node + v] = d;
            }
        }
    }

    // boundary extrapolation, Euler/space_time_predictor.f90:100-119
    for (int task = tt; task < 4 * n; task += nthr) {
        const int f = task / n, p = task - f * n;
        const double *cf = (f & 1) ? op->frc : op->flc;
        double Qb[V], Fb[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            double a = 0.0, c = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) {
                const int idx = (f < 2) ? (k + n * p) : (p + n * k);
                a = a + qhiL[v * NN + idx] * cf[k];
                c = c + (f < 2 ? fhxL[v * NN + idx] : fhyL[v * NN + idx]) * cf[k];
            }
            Qb[v] = a;
            Fb[v] = c;
        }
        if (act) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (TEST) {
                    t.qbnd[(size_t)cell * V * 4 * n + v + V * (f + 4 * p)] = Qb[v];
                    t.fbnd[(size_t)cell * V * 4 * n + v + V * (f + 4 * p)] = Fb[v];
                } else {
                    const size_t o = ((size_t)(f * V + v) * n + p) * b.G + g;
                    b.fq[o] = Qb[v];
                    b.ff[o] = Fb[v];
                }
            }
        }
    }
}

// CPW cells from local cell `first` of one solve (cell0, parity par); smem holds
// 2 * CPW * SLOT doubles.  TEST: cells of t.uh, outputs in the reference layouts.
template <class P, int N, bool TEST>
ADG_DEV void predictor_cells(const MeshArgs &m, const StepBufs &b, const TestBufs &t, double *smem, int first,
                             int cell0, int par, double dt)
{
    using S = Shape<P, N>;
    const double rdx = 1.0 / m.dx, rdy = 1.0 / m.dy;  // per solve; div_inv, gradients
    constexpr int n = S::n, NN = S::NN, CPW = S::CPW, V = S::V, SLOT = S::SLOT;
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const int cell = first + lc;                 // local cell of the solve
    const size_t g = (size_t)cell0 + cell;       // cell of the group arrays
    const bool own = lc < CPW;
    const bool act = own && cell < m.ncell;
    const int lcs = own ? lc : CPW - 1;
    auto SM = [&](int sl) { return smem + (sl * CPW + lcs) * SLOT; };

    double u[V];
    const double *uh = TEST ? t.uh : b.uh[par];
    const size_t gu = TEST ? (size_t)cell : g;
#pragma unroll
    for (int v = 0; v < V; v++) u[v] = act ? uh[gu * V * NN + V * tt + v] : 0.0;
    if (!act) {  // an admissible dummy state for idle lanes
        if (V == 4) {
            u[0] = 1.0;
            u[3] = 2.5;
        } else {
            u[0] = 0.5; u[3] = 2.5; u[4] = 0.5; u[7] = 2.5; u[8] = 0.5;
        }
    }
    const double wi = op->w[i], wj = op->w[j];
    double kx_i[n], kx_j[n], du_i[n], du_j[n];
#pragma unroll
    for (int k = 0; k < n; k++) {
        kx_i[k] = op->kxi[k + n * i];
        kx_j[k] = op->kxi[k + n * j];
        if (P::kNcp) {
            du_i[k] = op->dudx[i + n * k];
            du_j[k] = op->dudx[j + n * k];
        }
    }

    double q[V][n], fh[2][V], sh[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        fh[0][v] = 0.0;
        fh[1][v] = 0.0;
        sh[v] = 0.0;
#pragma unroll
        for (int l = 0; l < n; l++) q[v][l] = u[v];
    }
    int slot = 0;
    {
        // The first Picard sweep starts from the trivial initial guess (every time slice of
        // the node holds luh, Euler/space_time_predictor.f90:27-35): the fluxes, contractions
        // and NCP source of its slices are the same numbers, formed once here; the slices
        // differ only by their weights.
        double F[2 * V], G[2 * V], ax[V], ay[V], src[V];
        P::flux(m, u, F);
        double *buf = SM(0);
        if (own) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                buf[v * NN + j * n + i] = F[v];
                buf[V * NN + v * NN + i * n + j] = F[V + v];
                if (P::kNcp) buf[2 * V * NN + v * NN + tt] = u[v];
            }
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) acc = acc + buf[v * NN + j * n + k] * kx_i[k];
            ax[v] = acc;
            acc = 0.0;
#pragma unroll
            for (int k = 0; k < n; k++) acc = acc + buf[V * NN + v * NN + i * n + k] * kx_j[k];
            ay[v] = acc;
            if (P::kNcp) {  // gradients, Baer-Nunziato/space_time_predictor.f90:72,79
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + k + n * j] * du_i[k];
                G[v] = rdx * acc;
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + i + n * k] * du_j[k];
                G[V + v] = rdy * acc;
            }
        }
        if (P::kNcp) P::ncp(m, u, G, src);  // NCP source, Baer-Nunziato/space_time_predictor.f90:84-92
        double qn[V][n];
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
            for (int ll = 0; ll < n; ll++) qn[v][ll] = 0.0;
#pragma unroll
        for (int l = 0; l < n; l++) {
            const double wl = op->w[l];
            const double sx = div_inv(wl * wj * dt, m.dx, rdx), sy = div_inv(wl * wi * dt, m.dy, rdy);
            const double c0 = wi * wj * op->f0[l];
            const double f = wi * wj * wl * dt;
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double rr = c0 * u[v] - sx * ax[v];  // :66
                double rv = rr - sy * ay[v];                // :72
                if (P::kNcp) rv = rv + f * src[v];
                // time contraction, :79-83, accumulated slice by slice
#pragma unroll
                for (int ll = 0; ll < n; ll++) qn[v][ll] = qn[v][ll] + rv * op->ik1[ll + n * l];
            }
        }
        const double f1 = 1.0 / (wi * wj);  // :81
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
            for (int ll = 0; ll < n; ll++) q[v][ll] = f1 * qn[v][ll];
        slot = 1;  // the next exchange goes to the other buffer: SM(0) may still be read
    }
    // The remaining Picard iterations, Euler/space_time_predictor.f90:49-84
    // (Baer-Nunziato/space_time_predictor.f90:54-104).  Both loops stay rolled: the time line q
    // rotates so that the current slice is always q[:,0], and the time contraction with iK1
    // accumulates slice by slice (summation order over m unchanged), which keeps the register
    // footprint at two time lines.
#pragma unroll 1
    for (int it = 1; it <= N; it++) {
        double qn[V][n];
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
            for (int ll = 0; ll < n; ll++) qn[v][ll] = 0.0;
#pragma unroll 1
        for (int l = 0; l < n; l++) {
            double Q[V], F[2 * V];
#pragma unroll
            for (int v = 0; v < V; v++) Q[v] = q[v][0];
            P::flux(m, Q, F);
            double *buf = SM(slot);
            if (own) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    buf[v * NN + j * n + i] = F[v];
                    buf[V * NN + v * NN + i * n + j] = F[V + v];
                    if (P::kNcp) buf[2 * V * NN + v * NN + tt] = Q[v];
                }
            }
            const double wl = op->w[l];
            if (it == N) {  // time average of the last sweep's fluxes (:92)
#pragma unroll
                for (int v = 0; v < V; v++) {
                    fh[0][v] = fh[0][v] + F[v] * wl;
                    fh[1][v] = fh[1][v] + F[V + v] * wl;
                }
            }
            __syncthreads();
            const double sx = div_inv(wl * wj * dt, m.dx, rdx), sy = div_inv(wl * wi * dt, m.dy, rdy);
            const double c0 = wi * wj * op->f0[l];
            double G[2 * V], rv[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + buf[v * NN + j * n + k] * kx_i[k];
                const double rr = c0 * u[v] - sx * acc;  // :66
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + buf[V * NN + v * NN + i * n + k] * kx_j[k];
                rv[v] = rr - sy * acc;  // :72
                if (P::kNcp) {          // gradients, Baer-Nunziato/space_time_predictor.f90:72,79
                    acc = 0.0;
#pragma unroll
                    for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + k + n * j] * du_i[k];
                    G[v] = rdx * acc;
                    acc = 0.0;
#pragma unroll
                    for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + i + n * k] * du_j[k];
                    G[V + v] = rdy * acc;
                }
            }
            if (P::kNcp) {  // NCP source, Baer-Nunziato/space_time_predictor.f90:84-92
                double src[V];
                P::ncp(m, Q, G, src);
                const double f = wi * wj * wl * dt;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    rv[v] = rv[v] + f * src[v];
                    if (it == N) sh[v] = sh[v] + src[v] * wl;  // lShi (:111)
                }
            }
            // time contraction, :79-83: q_new(v,ll) = sum over m of rhs(v,m) iK1(ll,m), m = l here
#pragma unroll
            for (int v = 0; v < V; v++)
#pragma unroll
                for (int ll = 0; ll < n; ll++) qn[v][ll] = qn[v][ll] + rv[v] * op->ik1[ll + n * l];
            // rotate the time line: slice l+1 moves to position 0
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double q0 = q[v][0];
#pragma unroll
                for (int ll = 0; ll + 1 < n; ll++) q[v][ll] = q[v][ll + 1];
                q[v][n - 1] = q0;
            }
            slot ^= 1;
        }
        const double f = 1.0 / (wi * wj);  // :81
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
            for (int ll = 0; ll < n; ll++) q[v][ll] = f * qn[v][ll];
    }
    double qh[V];  // :90
#pragma unroll
    for (int v = 0; v < V; v++) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < n; l++) acc = acc + q[v][l] * op->w[l];
        qh[v] = acc;
    }
    __syncthreads();
    double *qhiL = SM(0);
    double *fhxL = SM(0) + V * NN;
    double *fhyL = SM(1);
    double *shL = SM(1) + V * NN;
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            qhiL[v * NN + tt] = qh[v];
            fhxL[v * NN + tt] = fh[0][v];
            fhyL[v * NN + tt] = fh[1][v];
            if (P::kNcp) shL[v * NN + tt] = sh[v];
        }
    }
    __syncthreads();
    predictor_finish<P, N, TEST>(m, b, t, qhiL, fhxL, fhyL, shL, tt, NN, act, cell, cell0);
}


// ---------------------------------------------------------------------------
// Lean time-line variant: the thread/node mapping of predictor_cells with a
// smaller register footprint.  The right-hand side of slice l replaces q(:, l)
// in the time line once the slice is done (after the sweep the time line holds
// rhs(:, 0..N) and the iK1 contraction runs in place, in the same summation
// order), the initial state and the operator rows live in LDS, the flux
// exchange is single-buffered (two barriers per slice) and the last sweep,
// which also accumulates the time averages, is a separate copy of the loop.
// smem: CPW cells x (V NN state + SLOT exchange) + the Kxi and dudx matrices.
// Lane k of this lane's quad, read through DPP (quad_perm [k,k,k,k]) on both halves of the
// double: a VALU move instead of an LDS round trip.  Every lane of the quad must be active.
template <int K>
ADG_DEV double quad_bcast(double x)
{
    constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, ctrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
// sum over k = 0..3, in that order, of (x of quad lane k) * c[k]
ADG_DEV double quad_dot4(double x, const double *c)
{
    double acc = 0.0;
    acc = acc + quad_bcast<0>(x) * c[0];
    acc = acc + quad_bcast<1>(x) * c[1];
    acc = acc + quad_bcast<2>(x) * c[2];
    acc = acc + quad_bcast<3>(x) * c[3];
    return acc;
}

// ---------------------------------------------------------------------------
// Matrix-core contractions (product build).  v_mfma_f64_4x4x4_4b_f64 multiplies four
// independent 4x4 blocks per wave; the block of lane L is (L >> 2) & 3, and within it
//   A[m][k] sits in lane 16 k + 4 blk + m,  B[k][n] in lane 16 k + 4 blk + n,
//   D[m][n] in lane 16 m + 4 blk + n
// (measured on gfx950 with one-hot operands).  A cell of 16 nodes is one block when lane
// 16 j + 4 blk + i holds node (i, j): then B = a nodal field X(i, j) contracts over j
// (D = Op . X, Op in A), and the lane-transposed field X(j, i) in A contracts over i
// (D = X^T . Op, Op in B), each landing in the node layout again.
ADG_DEV double mfma4(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, 0.0, 0, 0, 0);
#else
    return a * b;  // the host pass only type-checks device code
#endif
}
// x of lane addr / 4 of this wave (ds_bpermute on both halves: through the LDS crossbar, no
// LDS storage and no barrier)
ADG_DEV double lane_get(double x, int addr)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// ---------------------------------------------------------------------------
// LDS doubles of a cell are padded so that the cell's region starts NN doubles (mod 32) after
// the previous cell's: a lane's node-indexed element (tt) then lies 8 bytes after the previous
// lane's across cell boundaries, so any 32 consecutive lanes of a wave touch 32 distinct bank
// pairs (64 banks of 4 bytes).  At N=3 this is the pad of 16 that a sweep of even pads 0..24
// found best; at N=2 (nine lanes per cell) it is 5: the Baer-Nunziato N=2 predictor
// 1.248 -> 1.17 ms on 4 x 384^2 meshes, where even pads had left 2-way conflicts.
template <class P, int N>
struct LeanShape {
    using S = Shape<P, N>;
    static constexpr int BODY = S::V * S::NN + S::SLOT;
    static constexpr int PAD = ((S::NN - BODY) % 32 + 32) % 32;
    static constexpr int CELL = BODY + PAD;                // LDS doubles per cell
    static constexpr int LDS = S::CPW * CELL + 2 * S::NN;  // + Kxi, dudx
};

template <class P, int N, bool TEST, class M>
ADG_DEV void predictor_lean(const M &m, const StepBufs &b, const TestBufs &t, double *smem, int first,
                            int cell0, int par, double dt)
{
    using S = Shape<P, N>;
    const double rdx = 1.0 / m.dx, rdy = 1.0 / m.dy;  // per solve; div_inv, gradients
    using L = LeanShape<P, N>;
    constexpr int n = S::n, NN = S::NN, CPW = S::CPW, V = S::V;
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const int cell = first + lc;  // local cell of the solve
    const bool own = lc < CPW;
    const bool act = own && cell < m.ncell;
    double *sC = smem + (own ? lc : CPW - 1) * L::CELL;
    double *sU = sC;            // [v][tt] initial state
    double *buf = sC + V * NN;  // Fx [v][j][i], Fy [v][i][j], q [v][tt]
    double *sK = smem + CPW * L::CELL, *sD = sK + NN;
    for (int k = tid; k < NN; k += blockDim.x) {
        sK[k] = op->kxi[k];
        sD[k] = op->dudx[k];
    }
    {
        const double *uh = TEST ? t.uh : b.uh[par];
        const size_t gu = TEST ? (size_t)cell : (size_t)cell0 + cell;
        double u[V];
#pragma unroll
        for (int v = 0; v < V; v++) u[v] = act ? uh[gu * V * NN + V * tt + v] : 0.0;
        if (!act) {  // an admissible dummy state for idle lanes
            if (V == 4) {
                u[0] = 1.0;
                u[3] = 2.5;
            } else {
                u[0] = 0.5; u[3] = 2.5; u[4] = 0.5; u[7] = 2.5; u[8] = 0.5;
            }
        }
        if (own) {
#pragma unroll
            for (int v = 0; v < V; v++) sU[v * NN + tt] = u[v];
        }
    }
    __syncthreads();
    const double wi = op->w[i], wj = op->w[j];
    const double f1 = 1.0 / (wi * wj);  // :81
    // DPPX: at n = 4 the four nodes of a row (i fastest) are this lane's quad, so the x
    // contractions of the fluxes (and, Baer-Nunziato, of q for the gradients) take the row
    // values from the other lanes of the quad by DPP instead of an LDS round trip: the same
    // values in the same order, so the same bits.  Every lane of the workgroup owns a node.
    // (Baer-Nunziato N=3 predictor 2.73 -> 2.58 ms, Euler N=3 1.88 -> 1.75 ms on 4 x 384^2 /
    // 4 x 512^2 meshes.)
    constexpr bool DPPX = n == 4 && CPW * NN == kPredBlock<P>;
    double kx[4], gxr[4];
    if constexpr (DPPX) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            kx[k] = sK[k + n * i];
            gxr[k] = sD[i + n * k];
        }
    }
    // OPR (n <= 3): the operator rows of this lane's contractions held in registers for the
    // whole kernel (4 n doubles) instead of re-read from LDS at every slice: halves the LDS
    // reads of the contractions (Baer-Nunziato N=2 predictor -2%, with the pad above -7.5%)
    constexpr bool OPR = !DPPX && n <= 3;
    double okx[n], oky[n], ogx[n], ogy[n];
    if constexpr (OPR) {
#pragma unroll
        for (int k = 0; k < n; k++) {
            okx[k] = sK[k + n * i];
            oky[k] = sK[k + n * j];
            ogx[k] = sD[i + n * k];
            ogy[k] = sD[j + n * k];
        }
    }
    // operator entries of the x / y flux contractions and the x / y gradients
    auto opkx = [&](int k) { if constexpr (OPR) return okx[k]; else return sK[k + n * i]; };
    auto opky = [&](int k) { if constexpr (OPR) return oky[k]; else return sK[k + n * j]; };
    auto opgx = [&](int k) { if constexpr (OPR) return ogx[k]; else return sD[i + n * k]; };
    auto opgy = [&](int k) { if constexpr (OPR) return ogy[k]; else return sD[j + n * k]; };
    double q[V][n];
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
        for (int l = 0; l < n; l++) q[v][l] = sU[v * NN + tt];
    double fh[2][V], sh[V];

    // fluxes of state Q at this thread's node, exchanged with the cell's threads, and the
    // slice's x / y flux contractions ax, ay (and, Baer-Nunziato, the gradients G).  A
    // variable whose flux is identically zero (Baer-Nunziato's volume fraction) is neither
    // exchanged nor contracted: its contraction is a sum of zeros, +0, so it is set to +0,
    // the same bits.  Only the gradients the NCP source reads are formed.
    auto slice_terms = [&](const double *Q, double *F, double *ax, double *ay, double *G) {
        P::flux(m, Q, F);
        if (own) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (!P::flux_zero(v)) {
                    if (!DPPX) buf[v * NN + j * n + i] = F[v];
                    buf[V * NN + v * NN + i * n + j] = F[V + v];
                }
                if (P::kNcp && P::grad_used(v)) buf[2 * V * NN + v * NN + tt] = Q[v];
            }
        }
        if constexpr (DPPX) {
#pragma unroll
            for (int v = 0; v < V; v++) ax[v] = P::flux_zero(v) ? 0.0 : quad_dot4(F[v], kx);
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
            if constexpr (!DPPX) {
#pragma unroll
                for (int k = 0; k < n && !P::flux_zero(v); k++) acc = acc + buf[v * NN + j * n + k] * opkx(k);
                ax[v] = acc;
            }
            acc = 0.0;
#pragma unroll
            for (int k = 0; k < n && !P::flux_zero(v); k++) acc = acc + buf[V * NN + v * NN + i * n + k] * opky(k);
            ay[v] = acc;
            G[v] = G[V + v] = 0.0;
            if (P::kNcp && P::grad_used(v)) {  // gradients, Baer-Nunziato/space_time_predictor.f90:72,79
                acc = 0.0;
                if constexpr (DPPX) {
                    acc = quad_dot4(Q[v], gxr);
                } else {
#pragma unroll
                    for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + k + n * j] * opgx(k);
                }
                G[v] = rdx * acc;
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + i + n * k] * opgy(k);
                G[V + v] = rdy * acc;
            }
        }
        __syncthreads();  // the exchange buffer is free for the next slice
    };

    // one Picard sweep, Euler/space_time_predictor.f90:49-84 (Baer-Nunziato/space_time_predictor.f90:54-104).
    // FIRST: the first sweep starts from the trivial initial guess, every time slice of the
    // node holding luh (:27-35), so the fluxes, contractions and NCP source of its slices are
    // the same numbers: they are formed once and the slices differ only by their weights.
    auto sweep = [&](auto last_tag, auto first_tag) {
        constexpr bool LAST = decltype(last_tag)::value;
        constexpr bool FIRST = decltype(first_tag)::value;
        if (LAST) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                fh[0][v] = 0.0;
                fh[1][v] = 0.0;
                sh[v] = 0.0;
            }
        }
        if constexpr (FIRST) {
            double Q[V], F[2 * V], G[2 * V], ax0[V], ay0[V], src0[V];
#pragma unroll
            for (int v = 0; v < V; v++) Q[v] = q[v][0];
            slice_terms(Q, F, ax0, ay0, G);
            if (P::kNcp) P::ncp(m, Q, G, src0);  // NCP source, Baer-Nunziato/space_time_predictor.f90:84-92
            // the right-hand sides of the slices in place of the time line
#pragma unroll
            for (int l = 0; l < n; l++) {
                const double wl = op->w[l];
                const double sx = div_inv(wl * wj * dt, m.dx, rdx), sy = div_inv(wl * wi * dt, m.dy, rdy);
                const double c0 = wi * wj * op->f0[l];
                const double f = wi * wj * wl * dt;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    const double rr = c0 * sU[v * NN + tt] - sx * ax0[v];  // :66
                    double r = rr - sy * ay0[v];                           // :72
                    if (P::kNcp && !P::src_zero(v)) r = r + f * src0[v];
                    q[v][l] = r;
                }
            }
        } else {
#pragma unroll 1
        for (int l = 0; l < n; l++) {
            double Q[V], axd[V];
#pragma unroll
            for (int v = 0; v < V; v++) Q[v] = q[v][0];
            {
                double F[2 * V];
                P::flux(m, Q, F);
                if (own) {
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        if (!P::flux_zero(v)) {
                            if (!DPPX) buf[v * NN + j * n + i] = F[v];
                            buf[V * NN + v * NN + i * n + j] = F[V + v];
                        }
                        if (P::kNcp && P::grad_used(v)) buf[2 * V * NN + v * NN + tt] = Q[v];
                    }
                }
                if constexpr (DPPX) {
#pragma unroll
                    for (int v = 0; v < V; v++) axd[v] = P::flux_zero(v) ? 0.0 : quad_dot4(F[v], kx);
                }
                if (LAST) {  // time average of the last sweep's fluxes (:92)
                    const double wl = op->w[l];
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        if (!P::flux_zero(v)) fh[0][v] = fh[0][v] + F[v] * wl;
                        if (!P::flux_zero(v)) fh[1][v] = fh[1][v] + F[V + v] * wl;
                    }
                }
            }
            __syncthreads();
            const double wl = op->w[l];
            const double sx = div_inv(wl * wj * dt, m.dx, rdx), sy = div_inv(wl * wi * dt, m.dy, rdy);
            const double c0 = wi * wj * op->f0[l];
            double rv[V], G[2 * V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
                if constexpr (DPPX) {
                    acc = axd[v];
                } else {
#pragma unroll
                    for (int k = 0; k < n && !P::flux_zero(v); k++) acc = acc + buf[v * NN + j * n + k] * opkx(k);
                }
                const double rr = c0 * sU[v * NN + tt] - sx * acc;  // :66
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n && !P::flux_zero(v); k++) acc = acc + buf[V * NN + v * NN + i * n + k] * opky(k);
                rv[v] = rr - sy * acc;  // :72
                G[v] = G[V + v] = 0.0;
                if (P::kNcp && P::grad_used(v)) {  // gradients, Baer-Nunziato/space_time_predictor.f90:72,79
                    acc = 0.0;
                    if constexpr (DPPX) {
                        acc = quad_dot4(Q[v], gxr);
                    } else {
#pragma unroll
                        for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + k + n * j] * opgx(k);
                    }
                    G[v] = rdx * acc;
                    acc = 0.0;
#pragma unroll
                    for (int k = 0; k < n; k++) acc = acc + buf[2 * V * NN + v * NN + i + n * k] * opgy(k);
                    G[V + v] = rdy * acc;
                }
            }
            __syncthreads();  // the exchange buffer is free for the next slice
            if (P::kNcp) {  // NCP source, Baer-Nunziato/space_time_predictor.f90:84-92
                double src[V];
                P::ncp(m, Q, G, src);
                const double f = wi * wj * wl * dt;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    if (!P::src_zero(v)) rv[v] = rv[v] + f * src[v];
                    if (LAST && !P::src_zero(v)) sh[v] = sh[v] + src[v] * wl;  // lShi (:111)
                }
            }
            // slice l is done: its right-hand side takes its place, and the time line rotates
#pragma unroll
            for (int v = 0; v < V; v++) {
#pragma unroll
                for (int ll = 0; ll + 1 < n; ll++) q[v][ll] = q[v][ll + 1];
                q[v][n - 1] = rv[v];
            }
        }
        }
        // time contraction, :79-83: q_new(v,ll) = f1 * sum over m of rhs(v,m) iK1(ll,m)
#pragma unroll
        for (int v = 0; v < V; v++) {
            double qn[n];
#pragma unroll
            for (int ll = 0; ll < n; ll++) {
                double acc = 0.0;
#pragma unroll
                for (int mm = 0; mm < n; mm++) acc = acc + q[v][mm] * op->ik1[ll + n * mm];
                qn[ll] = acc;
            }
#pragma unroll
            for (int ll = 0; ll < n; ll++) q[v][ll] = f1 * qn[ll];
        }
    };
    sweep(std::false_type{}, std::true_type{});
#pragma unroll 1
    for (int it = 1; it < N; it++) sweep(std::false_type{}, std::false_type{});
    sweep(std::true_type{}, std::false_type{});

    double qh[V];  // :90
#pragma unroll
    for (int v = 0; v < V; v++) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < n; l++) acc = acc + q[v][l] * op->w[l];
        qh[v] = acc;
    }
    double *qhiL = buf, *fhxL = buf + V * NN, *fhyL = sU, *shL = P::kNcp ? buf + 2 * V * NN : sU;
    __syncthreads();
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            qhiL[v * NN + tt] = qh[v];
            fhxL[v * NN + tt] = fh[0][v];
            fhyL[v * NN + tt] = fh[1][v];
            if (P::kNcp) shL[v * NN + tt] = sh[v];
        }
    }
    __syncthreads();
    predictor_finish<P, N, TEST>(m, b, t, qhiL, fhxL, fhyL, shL, tt, NN, act, cell, cell0);
}

// ---------------------------------------------------------------------------
// Matrix-core predictor (product build, N = 2 and 3).  Thread mapping: lane 16 j + 4 b + i of
// wave w holds the time line of node (i, j) of local cell 4 w + b, so each cell is one block of
// v_mfma_f64_4x4x4_4b_f64 (mfma4 above).  At N = 2 the 3 x 3 nodes sit in a 4 x 4 block whose
// padding lanes carry an admissible dummy state that stays put (no flux terms) and meet only
// zero operator entries.  Per time slice the x / y flux contractions and the Baer-Nunziato
// gradients are matrix products (fields contracted over i are transposed across the cell's
// lanes by ds_bpermute); after the last sweep the volume integral and the boundary
// extrapolation are matrix products too, and every lane writes its own lduh node and its own
// (face, point) of fq / ff: nothing is exchanged through LDS and no barrier is needed.  LDS
// only holds the initial state: in the persistent kernel each wave's four cells, copied from
// uh by the asynchronous global-to-LDS loads of gfx950 while the wave works on its previous
// entry (two buffers, entries alternate), so no entry waits on HBM for its input.  The
// matrix core sums each 4-term contraction in its own order (measured: the order of an FMA
// chain over k), within the product build's parity.
// ---------------------------------------------------------------------------
// Non-temporal 16-byte loads and stores (streamed data that no other workgroup of the kernel
// reads back: their lines need not stay in L2, where they would push out shared lines)
typedef double ntd2 __attribute__((ext_vector_type(2)));
ADG_DEV double2 nt_load2(const double2 *p)
{
    const ntd2 x = __builtin_nontemporal_load((const ntd2 *)p);
    return make_double2(x.x, x.y);
}
ADG_DEV void nt_store2(double2 *p, double2 v)
{
    ntd2 x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, (ntd2 *)p);
}

// s_waitcnt immediate that waits for vmcnt(0) only (expcnt and lgkmcnt at their maxima)
constexpr int kWaitVmcnt0 = 0x0F70;
// the admissible state idle and padding lanes carry (the dummy of every predictor variant)
template <class P>
ADG_DEV void mf_dummy(double *u)
{
#pragma unroll
    for (int v = 0; v < P::V; v++) u[v] = 0.0;
    if (P::V == 4) {
        u[0] = 1.0;
        u[3] = 2.5;
    } else {
        u[0] = 0.5; u[3] = 2.5; u[4] = 0.5; u[7] = 2.5; u[8] = 0.5;
    }
}

// threads of the matrix-core predictor's workgroup (four cells per wave): four waves, so that a
// workgroup writes each face-extrapolation plane's 128-byte line of 16 cells at once (with one
// wave, Baer-Nunziato's two arenas ran 2-3 ms per ensemble step slower side by side)
template <class P>
constexpr int kMfBlock = 256;

template <class P, int N>
struct MfShape {
    static constexpr int n = N + 1, NN = n * n, V = P::V;
    static constexpr int CPW = kMfBlock<P> / 16;  // cells per workgroup: four per wave
    static constexpr int LDS = CPW * V * 16;        // test kernel: initial state per cell, [v][16 lanes]
    static constexpr int CELLS = 4 * V * NN;        // doubles of a wave's four cells, uh layout
    // bytes per lane of one global-to-LDS load: 16 where every cell starts 16-byte aligned
    static constexpr int CHUNK = (V * NN) % 2 == 0 ? 16 : 4;
    static constexpr int LOADS = (CELLS * 8 + 64 * CHUNK - 1) / (64 * CHUNK);  // per wave and entry
    static_assert(n <= 4, "a cell's nodes must fit one 4 x 4 block");
};

// Issue the asynchronous copy of the first `ncells` (<= 4) cells of a wave's entry from uh
// (src, the entry's first cell) into dst (LDS, the uh layout); completes with vmcnt.
template <class P, int N>
ADG_DEV void mf_prefetch(const double *src, double *dst, int ncells)
{
    using S = MfShape<P, N>;
    const int lane = threadIdx.x & 63;
    const int bytes = ncells * S::V * S::NN * 8;
    const char *s = (const char *)src;
    char *d = (char *)dst;
#pragma unroll
    for (int r = 0; r < S::LOADS; r++) {
        const int off = (r * 64 + lane) * S::CHUNK;
        if (off < bytes)
            __builtin_amdgcn_global_load_lds((const void *)(s + off),
                                             (__attribute__((address_space(3))) void *)(d + r * 64 * S::CHUNK),
                                             S::CHUNK, 0, 0);
    }
}

// D = A B + C of the four blocks (mfma4 with an accumulator)
ADG_DEV double mfma4c(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
#else
    return a * b + c;  // the host pass only type-checks device code
#endif
}

struct NoHook {
    ADG_DEV void operator()() const {}
};

// Operator entries the matrix-core predictor indexes per lane, staged in LDS by the persistent
// kernel (offsets in doubles): kxi, dudx (n x n), w, flc, frc (n), then the dummy state.
enum { kSopKxi = 0, kSopDudx = 16, kSopW = 32, kSopFlc = 36, kSopFrc = 40, kSopPad = 44 };
template <class P>
constexpr int kSopSize = kSopPad + P::V;

// PF: the initial state comes from ubuf, the wave's four cells in LDS (uh layout) whose
// asynchronous copy was issued earlier, and the per-lane operator entries from sop (LDS, the
// layout above); after() runs once the initial state is read (the caller's next prefetch).
// After the sweeps ubuf stages the cells' lduh, which the wave then stores as whole rows.
// Otherwise (test kernel) both come from global memory, the state staged in smem.
template <class P, int N, bool TEST, bool PF, class M, class After = NoHook>
ADG_DEV void predictor_mf(const M &m, const StepBufs &b, const TestBufs &t, double *smem, int first, int cell0,
                          int par, double dt, double *ubuf = nullptr, const double *sop = nullptr,
                          After after = {})
{
    using S = MfShape<P, N>;
    constexpr int n = S::n, NN = S::NN, V = S::V;
    const double rdx = 1.0 / m.dx, rdy = 1.0 / m.dy;  // per solve; div_inv, gradients
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lane = tid & 63;
    const int i = lane & 3, j = lane >> 4, lc = 4 * (tid >> 6) + ((lane >> 2) & 3);
    const bool real = i < n && j < n;  // a node of the cell, not padding
    const int tt = real ? i + n * j : 0;  // node index in the reference order
    const int cell = first + lc;  // local cell of the solve
    const bool act = real && cell < m.ncell;
    const bool cellok = cell < m.ncell;
    // ds_bpermute address of the lane holding node (j, i) of the same cell
    const int xaddr = ((i << 4) | (lane & 12) | j) << 2;
    auto kxi = [&](int k) { if constexpr (PF) return sop[kSopKxi + k]; else return op->kxi[k]; };
    auto dudx = [&](int k) { if constexpr (PF) return sop[kSopDudx + k]; else return op->dudx[k]; };
    auto wq = [&](int k) { if constexpr (PF) return sop[kSopW + k]; else return op->w[k]; };
    // operand entries of this lane (zero outside the cell's n x n block):
    //   mk  = kxi(j, i): B of the x flux contraction, A of the y flux contraction
    //   md  = dudx(i, j): B of the x gradient, A of the y gradient
    // (the volume integral's kxi(i, j) and the boundary coefficients are read at the end)
    double mk = 0.0, md = 0.0;
    if (real) {
        mk = kxi(j + n * i);
        md = dudx(i + n * j);
    }
    // padding lanes: unit weights and no flux terms, so their time line stays the dummy state
    const double wi = real ? wq(i) : 1.0, wj = real ? wq(j) : 1.0;
    const double f1 = 1.0 / (wi * wj);  // :81
    const double tx = real ? dt : 0.0;  // the time step in the flux and source terms
    // this lane's initial state, element v at sU[SS * v]
    constexpr int SS = PF ? 1 : 16;
    const double *sU;
    double q[V][n];
    if constexpr (PF) {
        // the prefetched cells; idle and padding lanes read the dummy state
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        sU = act ? ubuf + (((lane >> 2) & 3) * NN + tt) * V : sop + kSopPad;
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
            for (int l = 0; l < n; l++) q[v][l] = sU[v];
    } else {
        double *sw = smem + lc * V * 16 + (lane & 3) + 4 * (lane >> 4);  // [v][16] of this cell
        const double *uh = TEST ? t.uh : b.uh[par];
        const size_t gu = TEST ? (size_t)cell : (size_t)cell0 + cell;
        double u[V];
#pragma unroll
        for (int v = 0; v < V; v++) u[v] = act ? uh[gu * V * NN + V * tt + v] : 0.0;
        if (!act) mf_dummy<P>(u);  // idle and padding lanes
#pragma unroll
        for (int v = 0; v < V; v++) {
            sw[16 * v] = u[v];
#pragma unroll
            for (int l = 0; l < n; l++) q[v][l] = u[v];
        }
        sU = sw;
    }
    after();
    double fh[2][V], sh[V];

    // fluxes of state Q at this lane's node, the slice's flux contractions ax, ay and
    // (Baer-Nunziato) the gradients G, Euler/space_time_predictor.f90:60-72,
    // Baer-Nunziato/space_time_predictor.f90:66-80
    auto slice_terms = [&](const double *Q, double *F, double *ax, double *ay, double *G) {
        P::flux(m, Q, F);
#pragma unroll
        for (int v = 0; v < V; v++) {
            ax[v] = P::flux_zero(v) ? 0.0 : mfma4(lane_get(F[v], xaddr), mk);
            ay[v] = P::flux_zero(v) ? 0.0 : mfma4(mk, F[V + v]);
            G[v] = G[V + v] = 0.0;
            if (P::kNcp && P::grad_used(v)) {
                G[v] = rdx * mfma4(lane_get(Q[v], xaddr), md);
                G[V + v] = rdy * mfma4(md, Q[v]);
            }
        }
    };
    // one Picard sweep, Euler/space_time_predictor.f90:49-84 (Baer-Nunziato/space_time_predictor.f90:54-104);
    // FIRST: from the trivial initial guess every slice holds luh (:27-35), so the slices share
    // their flux terms, formed once, and differ only by their weights
    auto sweep = [&](auto last_tag, auto first_tag) {
        constexpr bool LAST = decltype(last_tag)::value;
        constexpr bool FIRST = decltype(first_tag)::value;
        if (LAST) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                fh[0][v] = 0.0;
                fh[1][v] = 0.0;
                sh[v] = 0.0;
            }
        }
        if constexpr (FIRST) {
            double Q[V], F[2 * V], G[2 * V], ax[V], ay[V], src[V];
#pragma unroll
            for (int v = 0; v < V; v++) Q[v] = q[v][0];
            slice_terms(Q, F, ax, ay, G);
            if (P::kNcp) P::ncp(m, Q, G, src);  // :84-92
#pragma unroll
            for (int l = 0; l < n; l++) {
                const double wl = op->w[l];
                const double sx = div_inv(wl * wj * tx, m.dx, rdx), sy = div_inv(wl * wi * tx, m.dy, rdy);
                const double c0 = wi * wj * op->f0[l];
                const double f = wi * wj * wl * tx;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double r = (c0 * sU[SS * v] - sx * ax[v]) - sy * ay[v];  // :66, :72
                    if (P::kNcp && !P::src_zero(v)) r = r + f * src[v];
                    q[v][l] = r;
                }
            }
        } else {
            // two slices per trip where n is even: half the time line's rotation moves (BN N=3
            // 15.2 -> 14.6 ms; with three slices the remainder loop spilled, N=2 11.9 -> 28 ms)
            constexpr int UNR = n % 2 == 0 ? 2 : 1;
#pragma unroll UNR
            for (int l = 0; l < n; l++) {
                double Q[V], F[2 * V], ax[V], ay[V], G[2 * V], rv[V];
#pragma unroll
                for (int v = 0; v < V; v++) Q[v] = q[v][0];
                slice_terms(Q, F, ax, ay, G);
                const double wl = op->w[l];
                if (LAST) {  // time average of the last sweep's fluxes (:92)
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        if (!P::flux_zero(v)) fh[0][v] = fh[0][v] + F[v] * wl;
                        if (!P::flux_zero(v)) fh[1][v] = fh[1][v] + F[V + v] * wl;
                    }
                }
                const double sx = div_inv(wl * wj * tx, m.dx, rdx), sy = div_inv(wl * wi * tx, m.dy, rdy);
                const double c0 = wi * wj * op->f0[l];
#pragma unroll
                for (int v = 0; v < V; v++) rv[v] = (c0 * sU[SS * v] - sx * ax[v]) - sy * ay[v];  // :66, :72
                if (P::kNcp) {  // NCP source, :84-92
                    double src[V];
                    P::ncp(m, Q, G, src);
                    const double f = wi * wj * wl * tx;
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        if (!P::src_zero(v)) rv[v] = rv[v] + f * src[v];
                        if (LAST && !P::src_zero(v)) sh[v] = sh[v] + src[v] * wl;  // lShi (:111)
                    }
                }
                // slice l is done: its right-hand side takes its place, the time line rotates
#pragma unroll
                for (int v = 0; v < V; v++) {
#pragma unroll
                    for (int ll = 0; ll + 1 < n; ll++) q[v][ll] = q[v][ll + 1];
                    q[v][n - 1] = rv[v];
                }
            }
        }
        // time contraction, :79-83: q_new(v,ll) = f1 * sum over m of rhs(v,m) iK1(ll,m)
#pragma unroll
        for (int v = 0; v < V; v++) {
            double qn[n];
#pragma unroll
            for (int ll = 0; ll < n; ll++) {
                double acc = 0.0;
#pragma unroll
                for (int mm = 0; mm < n; mm++) acc = acc + q[v][mm] * op->ik1[ll + n * mm];
                qn[ll] = acc;
            }
#pragma unroll
            for (int ll = 0; ll < n; ll++) q[v][ll] = f1 * qn[ll];
        }
    };
    sweep(std::false_type{}, std::true_type{});
#pragma unroll 1
    for (int it = 1; it < N; it++) sweep(std::false_type{}, std::false_type{});
    sweep(std::true_type{}, std::false_type{});

    // time averages (:90), volume integral (Euler/volume_integral.f90:17-28, Baer-Nunziato/
    // volume_integral.f90:36-40) and boundary extrapolation (Euler/space_time_predictor.f90:100-119):
    // lane (i, j) keeps node (i, j) of lduh and point j of face i (left, right, bottom, top)
    const size_t g = (size_t)cell0 + cell;
    const int fc = i, pt = j;
    const bool facelane = cellok && pt < n;
    // STAGEF: the face extrapolations of the workgroup's 16 cells are staged in the entry's LDS
    // buffers and stored a 16-cell line of each plane at a time (each wave would write 32 bytes
    // of every 128-byte line, which L2 fetches before merging the four waves' parts)
    constexpr int PL = 4 * V * n;  // planes [face][v][p] of fq and of ff
    constexpr bool STAGEF = PF && !TEST && PL * 16 <= (kMfBlock<P> / 64) * S::CELLS;
    double qbv[STAGEF ? V : 1], fbv[STAGEF ? V : 1];
    //   mkt = kxi(i, j): B of the volume integral's x part, A of its y part
    //   mlr / mbt: flc / frc (j) in columns i = 0, 1 / i = 2, 3 (B of the boundary extrapolation)
    double mkt = 0.0, mlr = 0.0, mbt = 0.0;
    if (real) mkt = kxi(i + n * j);
    if (j < n) {
        double cf;
        if constexpr (PF)
            cf = sop[((i & 1) ? kSopFrc : kSopFlc) + j];
        else
            cf = (i & 1) ? op->frc[j] : op->flc[j];
        mlr = i < 2 ? cf : 0.0;
        mbt = i < 2 ? 0.0 : cf;
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
        double qh = 0.0;
#pragma unroll
        for (int l = 0; l < n; l++) qh = qh + q[v][l] * op->w[l];
        const double fxt = lane_get(fh[0][v], xaddr), qht = lane_get(qh, xaddr);
        const double qb = mfma4c(qh, mbt, mfma4(qht, mlr));
        const double fb = mfma4c(fh[1][v], mbt, mfma4(fxt, mlr));
        if (TEST) {
            if (act) {
                t.qhi[(size_t)cell * V * NN + v + V * tt] = qh;
                t.fhi[(size_t)cell * 2 * V * NN + v + V * (0 + 2 * tt)] = fh[0][v];
                t.fhi[(size_t)cell * 2 * V * NN + v + V * (1 + 2 * tt)] = fh[1][v];
                if (P::kNcp) t.shi[(size_t)cell * V * NN + v + V * tt] = sh[v];
            }
            if (facelane) {
                t.qbnd[(size_t)cell * V * 4 * n + v + V * (fc + 4 * pt)] = qb;
                t.fbnd[(size_t)cell * V * 4 * n + v + V * (fc + 4 * pt)] = fb;
            }
        } else {
            double d = 0.0 + mfma4(fxt, mkt) * wj / m.dx;
            d = d + mfma4(mkt, fh[1][v]) * wi / m.dy;
            if (P::kNcp) d = d + wi * wj * sh[v];
            if constexpr (PF) {
                if (act) ubuf[(((lane >> 2) & 3) * NN + tt) * V + v] = d;
            } else {
                if (act) b.lduh[g * V * NN + V * tt + v] = d;
            }
            if constexpr (STAGEF) {
                qbv[v] = qb;
                fbv[v] = fb;
            } else if (facelane) {
                const size_t o = ((size_t)(fc * V + v) * n + pt) * b.G + g;
                b.fq[o] = qb;
                b.ff[o] = fb;
            }
        }
    }
    if constexpr (PF && !TEST) {
        // the wave's lduh rows from LDS: consecutive lanes store consecutive 16-byte pairs, so
        // every 128-byte line is written whole by one instruction (single 8-byte stores of each
        // node would make L2 fetch every line before merging them).  A wave's LDS writes and
        // reads complete in order; the barrier only keeps the compiler from reordering them.
        __builtin_amdgcn_wave_barrier();
        const int c0w = first + 4 * (tid >> 6), nv = m.ncell - c0w < 4 ? m.ncell - c0w : 4;
        const int cnt = nv > 0 ? nv * V * NN : 0;
        double *dst = b.lduh + ((size_t)cell0 + c0w) * V * NN;
        if (((uintptr_t)dst & 15) == 0) {
            const double2 *s2 = (const double2 *)ubuf;
            double2 *d2 = (double2 *)dst;
            for (int k = lane; k < cnt / 2; k += 64) d2[k] = s2[k];
            if ((cnt & 1) && lane == 0) dst[cnt - 1] = ubuf[cnt - 1];
        } else {
            for (int k = lane; k < cnt; k += 64) dst[k] = ubuf[k];
        }
    }
    if constexpr (STAGEF) {
        __syncthreads();  // every wave has read its lduh rows: the entry's buffers take the planes
        double *stage = ubuf - (tid >> 6) * S::CELLS;  // the four waves' buffers of this entry
        const int nvw = m.ncell - first < 16 ? m.ncell - first : 16;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            double *out = pass ? b.ff : b.fq;
            if (facelane) {
#pragma unroll
                for (int v = 0; v < V; v++) stage[((fc * V + v) * n + pt) * 16 + lc] = pass ? fbv[v] : qbv[v];
            }
            __syncthreads();
            for (int k = tid; k < PL * 8; k += kMfBlock<P>) {
                const int plane = k >> 3, c2 = 2 * (k & 7);
                if (c2 >= nvw) continue;
                double *o = out + (size_t)plane * b.G + cell0 + first + c2;
                const double a0 = stage[plane * 16 + c2], a1 = stage[plane * 16 + c2 + 1];
                if (c2 + 1 < nvw && ((uintptr_t)o & 15) == 0) {
                    *(double2 *)o = make_double2(a0, a1);
                } else {
                    o[0] = a0;
                    if (c2 + 1 < nvw) o[1] = a1;
                }
            }
            __syncthreads();  // the stage is read before it is rewritten (or prefetched into)
        }
    }
}

// ---------------------------------------------------------------------------
// Space-time-node variant of the predictor: thread (cell, i, j, l) owns the
// state q(:, i, j, l) of one space-time node, so a cell's whole Picard sweep
// runs at once on (N+1)^3 threads and each thread holds one node instead of a
// time line (far fewer registers, more waves per SIMD).  Per sweep the x- and
// y-contractions (and the Baer-Nunziato gradients) read the other nodes' F
// (and q) of the same slice from LDS, and the time contraction reads the
// right-hand sides rhs(:, i, j, 0..N) of the node's time line.  Every sum is
// taken in the order of predictor_cells, so both variants give the same bits.
// BLK threads per workgroup hold BLK / (N+1)^3 cells; smem holds
// 2 V (N+1)^3 doubles per cell.
// ---------------------------------------------------------------------------
template <class P, int N, int BLK>
struct NodeShape {
    static constexpr int n = N + 1, NN = n * n, N3 = NN * n, V = P::V;
    static constexpr int CPB = BLK / N3;      // cells per workgroup
    static constexpr int R = 2 * V * N3;      // LDS doubles per cell
    static_assert(CPB >= 1, "a workgroup must hold a whole cell");
};

template <class P, int N, int BLK, bool TEST>
ADG_DEV void predictor_nodes(const MeshArgs &m, const StepBufs &b, const TestBufs &t, double *smem, int first,
                             int cell0, int par, double dt)
{
    using S = NodeShape<P, N, BLK>;
    const double rdx = 1.0 / m.dx, rdy = 1.0 / m.dy;  // per solve; div_inv, gradients
    constexpr int n = S::n, NN = S::NN, N3 = S::N3, V = S::V, CPB = S::CPB, R = S::R;
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / N3, tt = tid - lc * N3;
    const int ij = tt % NN, i = ij % n, j = ij / n, l = tt / NN;
    const int cell = first + lc;  // local cell of the solve
    const bool own = lc < CPB;
    const bool act = own && cell < m.ncell;
    double *sR = smem + (own ? lc : CPB - 1) * R;

    double u[V];
    const double *uh = TEST ? t.uh : b.uh[par];
    const size_t gu = TEST ? (size_t)cell : (size_t)cell0 + cell;
#pragma unroll
    for (int v = 0; v < V; v++) u[v] = act ? uh[gu * V * NN + V * ij + v] : 0.0;
    if (!act) {  // an admissible dummy state for idle lanes
        if (V == 4) {
            u[0] = 1.0;
            u[3] = 2.5;
        } else {
            u[0] = 0.5; u[3] = 2.5; u[4] = 0.5; u[7] = 2.5; u[8] = 0.5;
        }
    }
    const double wi = op->w[i], wj = op->w[j], wl = op->w[l];
    double kx_i[n], kx_j[n], du_i[n], du_j[n], ik_l[n];
#pragma unroll
    for (int k = 0; k < n; k++) {
        kx_i[k] = op->kxi[k + n * i];
        kx_j[k] = op->kxi[k + n * j];
        ik_l[k] = op->ik1[l + n * k];
        if (P::kNcp) {
            du_i[k] = op->dudx[i + n * k];
            du_j[k] = op->dudx[j + n * k];
        }
    }
    // the same expressions as predictor_cells for slice l of node (i, j)
    const double sx = div_inv(wl * wj * dt, m.dx, rdx), sy = div_inv(wl * wi * dt, m.dy, rdy);
    const double c0 = wi * wj * op->f0[l];
    const double fs = wi * wj * wl * dt;
    const double f1 = 1.0 / (wi * wj);

    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = u[v];
    double fhx[V], fhy[V], sh[V];  // time averages, valid in the l == 0 threads after the last sweep

    // one Picard sweep, Euler/space_time_predictor.f90:49-84 (Baer-Nunziato/space_time_predictor.f90:54-104)
    auto sweep = [&](auto last_tag) {
        constexpr bool LAST = decltype(last_tag)::value;
        double F[2 * V], rv[V], G[2 * V];
        P::flux(m, q, F);
        if (P::kNcp) {
            // phase 1: q and Fx of the slice
            if (own) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    sR[v * N3 + tt] = q[v];
                    sR[(V + v) * N3 + tt] = F[v];
                }
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[(V + v) * N3 + k + n * j + NN * l] * kx_i[k];
                rv[v] = c0 * u[v] - sx * acc;  // :66
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[v * N3 + k + n * j + NN * l] * du_i[k];
                G[v] = rdx * acc;
                if (LAST && l == 0) {  // time average of the last sweep's fluxes (:92)
                    double a = 0.0;
#pragma unroll
                    for (int mm = 0; mm < n; mm++) a = a + sR[(V + v) * N3 + ij + NN * mm] * op->w[mm];
                    fhx[v] = a;
                }
            }
            __syncthreads();
            // phase 2: Fy of the slice (q stays)
            if (own) {
#pragma unroll
                for (int v = 0; v < V; v++) sR[(V + v) * N3 + tt] = F[V + v];
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[(V + v) * N3 + i + n * k + NN * l] * kx_j[k];
                rv[v] = rv[v] - sy * acc;  // :72
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[v * N3 + i + n * k + NN * l] * du_j[k];
                G[V + v] = rdy * acc;
                if (LAST && l == 0) {
                    double a = 0.0;
#pragma unroll
                    for (int mm = 0; mm < n; mm++) a = a + sR[(V + v) * N3 + ij + NN * mm] * op->w[mm];
                    fhy[v] = a;
                }
            }
        } else {
            // Fx and Fy of the slice at once
            if (own) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    sR[v * N3 + tt] = F[v];
                    sR[(V + v) * N3 + tt] = F[V + v];
                }
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[v * N3 + k + n * j + NN * l] * kx_i[k];
                const double rr = c0 * u[v] - sx * acc;  // :66
                acc = 0.0;
#pragma unroll
                for (int k = 0; k < n; k++) acc = acc + sR[(V + v) * N3 + i + n * k + NN * l] * kx_j[k];
                rv[v] = rr - sy * acc;  // :72
                if (LAST && l == 0) {
                    double a = 0.0, c = 0.0;
#pragma unroll
                    for (int mm = 0; mm < n; mm++) a = a + sR[v * N3 + ij + NN * mm] * op->w[mm];
#pragma unroll
                    for (int mm = 0; mm < n; mm++) c = c + sR[(V + v) * N3 + ij + NN * mm] * op->w[mm];
                    fhx[v] = a;
                    fhy[v] = c;
                }
            }
        }
        double src[V];
        if (P::kNcp) {  // NCP source, Baer-Nunziato/space_time_predictor.f90:84-92
            P::ncp(m, q, G, src);
#pragma unroll
            for (int v = 0; v < V; v++) rv[v] = rv[v] + fs * src[v];
        }
        __syncthreads();
        // phase 3: time contraction, :79-83, q(v,ll) = f1 * sum over m of rhs(v,m) iK1(ll,m)
        if (own) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                sR[v * N3 + tt] = rv[v];
                if (P::kNcp && LAST) sR[(V + v) * N3 + tt] = src[v];
            }
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int mm = 0; mm < n; mm++) acc = acc + sR[v * N3 + ij + NN * mm] * ik_l[mm];
            q[v] = f1 * acc;
            if (P::kNcp && LAST && l == 0) {  // lShi (:111)
                double a = 0.0;
#pragma unroll
                for (int mm = 0; mm < n; mm++) a = a + sR[(V + v) * N3 + ij + NN * mm] * op->w[mm];
                sh[v] = a;
            }
        }
        __syncthreads();
    };
#pragma unroll 1
    for (int it = 0; it < N; it++) sweep(std::false_type{});
    sweep(std::true_type{});

    // qhi = sum over l of q(l) w(l), :90
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) sR[v * N3 + tt] = q[v];
    }
    __syncthreads();
    double qh[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        double acc = 0.0;
#pragma unroll
        for (int mm = 0; mm < n; mm++) acc = acc + sR[v * N3 + ij + NN * mm] * op->w[mm];
        qh[v] = acc;
    }
    __syncthreads();
    double *qhiL = sR, *fhxL = sR + V * NN, *fhyL = sR + 2 * V * NN, *shL = sR + 3 * V * NN;
    if (own && l == 0) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            qhiL[v * NN + ij] = qh[v];
            fhxL[v * NN + ij] = fhx[v];
            fhyL[v * NN + ij] = fhy[v];
            if (P::kNcp) shL[v * NN + ij] = sh[v];
        }
    }
    __syncthreads();
    predictor_finish<P, N, TEST>(m, b, t, qhiL, fhxL, fhyL, shL, tt, N3, act, cell, cell0);
}

template <class P, int N>
__global__ __launch_bounds__(kPredBlock<P>) void k_predictor(StepBufs b)
{
    using S = Shape<P, N>;
    __shared__ double smem[2 * S::CPW * S::SLOT];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.counters[kCntList] = 0;
        b.counters[kCntTroubled] = 0;
    }
    const int ntab = ld_const(b.meta + kTabPred);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabPred] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const SolveArgs sa = ld_const(b.solves + w.x);
        predictor_cells<P, N, false>(sa.m, b, TestBufs{}, smem, w.y, sa.cell0, st.parity, st.dt);
        __syncthreads();
    }
}

template <class P, int N>
__global__ __launch_bounds__(kPredBlock<P>) void k_predictor_test(MeshArgs m, StepBufs b, TestBufs t, double dt)
{
    using S = Shape<P, N>;
    __shared__ double smem[2 * S::CPW * S::SLOT];
    predictor_cells<P, N, true>(m, b, t, smem, blockIdx.x * S::CPW, 0, 0, dt);
}


// at least two waves per SIMD: the lean layout is meant to fit 256 registers
template <class P, int N>
__global__ __launch_bounds__(kPredBlock<P>) __attribute__((amdgpu_waves_per_eu(2))) void k_predictor_lean(StepBufs b)
{
    __shared__ double smem[LeanShape<P, N>::LDS];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.counters[kCntList] = 0;
        b.counters[kCntTroubled] = 0;
    }
    const int ntab = ld_const(b.meta + kTabPred);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabPred] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        // the solve's constants stay in the constant address space, read where used
        const ADG_CONST SolveArgs *sa = (const ADG_CONST SolveArgs *)(b.solves + w.x);
        predictor_lean<P, N, false>(sa->m, b, TestBufs{}, smem, w.y, sa->cell0, st.parity, st.dt);
        __syncthreads();
    }
}

template <class P, int N>
__global__ __launch_bounds__(kPredBlock<P>) void k_predictor_lean_test(MeshArgs m, StepBufs b, TestBufs t, double dt)
{
    __shared__ double smem[LeanShape<P, N>::LDS];
    predictor_lean<P, N, true>(m, b, t, smem, blockIdx.x * Shape<P, N>::CPW, 0, 0, dt);
}

// the matrix-core predictor (N <= 3, product build)
template <class P, int N>
__global__ __launch_bounds__(kMfBlock<P>) __attribute__((amdgpu_waves_per_eu(2))) void k_predictor_mfma(StepBufs b)
{
    using S = MfShape<P, N>;
    constexpr int WAVES = kMfBlock<P> / 64;
    // two buffers of each wave's four cells, alternating between the entries of the loop (a wave
    // only touches its own: no barrier in the loop), then the operator entries and dummy state
    __shared__ __align__(16) double lds[2 * WAVES * S::CELLS + kSopSize<P>];
    double *sop = lds + 2 * WAVES * S::CELLS;
    {
        constexpr int n = S::n;
        const OpPtr op = (OpPtr)b.op;
        const int k = threadIdx.x;
        if (k < n * n) {
            sop[kSopKxi + k] = op->kxi[k];
            sop[kSopDudx + k] = op->dudx[k];
        }
        if (k < n) {
            sop[kSopW + k] = op->w[k];
            sop[kSopFlc + k] = op->flc[k];
            sop[kSopFrc + k] = op->frc[k];
        }
        if (k == 0) {
            double u[P::V];
            mf_dummy<P>(u);
#pragma unroll
            for (int v = 0; v < P::V; v++) sop[kSopPad + v] = u[v];
        }
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.counters[kCntList] = 0;
        b.counters[kCntTroubled] = 0;
    }
    const int wv = threadIdx.x >> 6;
    const int ntab = ld_const(b.meta + kTabPred);
    // entry e's table row and solve state, read one entry ahead
    int2 w = make_int2(0, 0);
    SolveState st{};
    auto fetch = [&](int e) {
        w = ld_const(b.tab[kTabPred] + e);
        st = ld_const(b.state + w.x);
    };
    // the copy of this wave's four cells of the fetched entry into buffer k
    auto prefetch = [&](int k) {
        if (!st.active) return;
        const ADG_CONST SolveArgs *sa = (const ADG_CONST SolveArgs *)(b.solves + w.x);
        const int c = w.y + 4 * wv, nc = sa->m.ncell - c;
        if (nc > 0)
            mf_prefetch<P, N>(b.uh[st.parity] + ((size_t)sa->cell0 + c) * P::V * S::NN,
                              lds + (k * WAVES + wv) * S::CELLS, nc < 4 ? nc : 4);
    };
    int k = 0;
    int e = xcd_entry(blockIdx.x, gridDim.x);
    if (e < ntab) {
        fetch(e);
        prefetch(0);
    }
    for (; e < ntab; e += gridDim.x, k ^= 1) {
        const int2 wc = w;
        const SolveState sc = st;
        // the next entry's cells go to the other buffer, whose last reader was the previous entry;
        // issued once this entry's copy has landed (the wait for it counts every load before it)
        const int en = e + (int)gridDim.x;
        if (en < ntab) fetch(en);
        auto next = [&]() {
            if (en < ntab) prefetch(k ^ 1);
        };
        if (!sc.active) {
            next();
            continue;
        }
        const ADG_CONST SolveArgs *sa = (const ADG_CONST SolveArgs *)(b.solves + wc.x);
        predictor_mf<P, N, false, true>(sa->m, b, TestBufs{}, nullptr, wc.y, sa->cell0, sc.parity, sc.dt,
                                        lds + (k * WAVES + wv) * S::CELLS, sop, next);
    }
}

template <class P, int N>
__global__ __launch_bounds__(kMfBlock<P>) void k_predictor_mfma_test(MeshArgs m, StepBufs b, TestBufs t, double dt)
{
    __shared__ double smem[MfShape<P, N>::LDS];
    predictor_mf<P, N, true, false>(m, b, t, smem, blockIdx.x * MfShape<P, N>::CPW, 0, 0, dt);
}

template <class P, int N, int BLK>
__global__ __launch_bounds__(BLK) void k_predictor_nodes(StepBufs b)
{
    using S = NodeShape<P, N, BLK>;
    __shared__ double smem[S::CPB * S::R];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.counters[kCntList] = 0;
        b.counters[kCntTroubled] = 0;
    }
    const int ntab = ld_const(b.meta + kTabPred);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabPred] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const SolveArgs sa = ld_const(b.solves + w.x);
        predictor_nodes<P, N, BLK, false>(sa.m, b, TestBufs{}, smem, w.y, sa.cell0, st.parity, st.dt);
        __syncthreads();
    }
}

template <class P, int N, int BLK>
__global__ __launch_bounds__(BLK) void k_predictor_nodes_test(MeshArgs m, StepBufs b, TestBufs t, double dt)
{
    using S = NodeShape<P, N, BLK>;
    __shared__ double smem[S::CPB * S::R];
    predictor_nodes<P, N, BLK, true>(m, b, t, smem, blockIdx.x * S::CPB, 0, 0, dt);
}

// ---------------------------------------------------------------------------
// Face fluxes (solve_riemann_problem, Euler/riemann_problem.f90:10-116 and
// Baer-Nunziato/riemann_problem.f90:9-80): one thread per face point, reading
// the structure-of-arrays face extrapolations of the two neighbouring cells
// with consecutive lanes on consecutive cells, writing the flux (Euler) or
// the fluctuations D-/D+ (Baer-Nunziato) in the same order.
// ---------------------------------------------------------------------------
template <class P, int N>
__global__ __launch_bounds__(kBlock) void k_face(StepBufs b)
{
    constexpr int V = P::V, n = N + 1;
    const int ntab = ld_const(b.meta + kTabFace);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabFace] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + w.x);
        const auto &m = sa.m;
        const int I = m.imax, J = m.jmax;
        const long nfx = (long)(I + 1) * J, nfy = (long)I * (J + 1);
        const long t = (long)w.y + threadIdx.x;
        if (t >= (nfx + nfy) * n) continue;
        const double time = st.time;
        const bool dmr = V == 4 && m.bc_mode == kBcDmr;
        const size_t c0 = sa.cell0;
        const bool xdir = t < nfx * n;
        const long t2 = xdir ? t : t - nfx * n;
        const long nf = xdir ? nfx : nfy;
        const int p = (int)(t2 / nf);
        const long fidx = t2 - (long)p * nf;
        double QL[V], FL[V], QR[V], FR[V], Dm[V], Dp[V];
        auto ld = [&](int f, long c, double *Q, double *F) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                const size_t o = ((size_t)(f * V + v) * n + p) * b.G + c0 + c;
                Q[v] = b.fq[o];
                F[v] = b.ff[o];
            }
        };
        auto copy = [&](const double *Qs, const double *Fs, double *Qd, double *Fd) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                Qd[v] = Qs[v];
                Fd[v] = Fs[v];
            }
        };
        if (xdir) {
            const int fi = (int)(fidx % (I + 1)), j = (int)(fidx / (I + 1));
            if (fi > 0) ld(1, (fi - 1) + (long)I * j, QL, FL);
            if (fi < I) ld(0, fi + (long)I * j, QR, FR);
            if (fi == 0) {
                if (!dmr && m.side[0] == kSideWrap) {
                    ld(1, (I - 1) + (long)I * j, QL, FL);
                } else if (!dmr) {
                    side_ghost<P>(m, m.side[0], 1.0, 0.0, QR, FR, QL, FL);
                } else {  // DMR inflow, Euler/riemann_problem.f90:186-193
                    double Qbc[V], Fbc[2 * V];
#pragma unroll
                    for (int v = 0; v < V; v++) Qbc[v] = m.qa[v % 4];
                    P::flux(m, Qbc, Fbc);
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        QL[v] = Qbc[v];
                        FL[v] = 0.0 + Fbc[v] * 1.0 + Fbc[V + v] * 0.0;
                    }
                }
            }
            if (fi == I) {
                if (!dmr && m.side[1] == kSideWrap)
                    ld(0, (long)I * j, QR, FR);
                else if (!dmr)
                    side_ghost<P>(m, m.side[1], 1.0, 0.0, QL, FL, QR, FR);
                else
                    copy(QL, FL, QR, FR);  // the DMR right boundary is transmissive (:195-198)
            }
            P::face_flux(m, QL, FL, QR, FR, 1.0, 0.0, Dm, Dp);
        } else {
            const int i = (int)(fidx % I), fj = (int)(fidx / I);
            if (fj > 0) ld(3, i + (long)I * (fj - 1), QL, FL);
            if (fj < J) ld(2, i + (long)I * fj, QR, FR);
            if (fj == 0) {
                if (!dmr && m.side[2] == kSideWrap) {
                    ld(3, i + (long)I * (J - 1), QL, FL);
                } else if (!dmr) {
                    side_ghost<P>(m, m.side[2], 0.0, 1.0, QR, FR, QL, FL);
                } else {
                    copy(QR, FR, QL, FL);
                    if (dmr && !((double)(i + 1) * m.dx <= 1.0 / 6.0)) {  // reflective, :225-228
                        QL[2] = -QL[2];
                        FL[0] = -FL[0];
                        FL[1] = -FL[1];
                        FL[3] = -FL[3];
                    }
                }
            }
            if (fj == J) {
                if (!dmr && m.side[3] == kSideWrap) {
                    ld(2, i, QR, FR);
                } else if (!dmr) {
                    side_ghost<P>(m, m.side[3], 0.0, 1.0, QL, FL, QR, FR);
                } else {  // DMR top: exact shock position, :233-245
                    const bool pre = (double)(i + 1) * m.dx <= dmr_shock_x(m, time);
                    double Qbc[V], Fbc[2 * V];
#pragma unroll
                    for (int v = 0; v < V; v++) Qbc[v] = pre ? m.qa[v % 4] : m.qb[v % 4];
                    P::flux(m, Qbc, Fbc);
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        QR[v] = Qbc[v];
                        FR[v] = 0.0 + Fbc[v] * 0.0 + Fbc[V + v] * 1.0;
                    }
                }
            }
            P::face_flux(m, QL, FL, QR, FR, 0.0, 1.0, Dm, Dp);
        }
        double *out = xdir ? b.xflux : b.yflux;
        const size_t stride = xdir ? b.GX : b.GY;
        const size_t f0 = (size_t)(xdir ? sa.fx0 : sa.fy0) + fidx;
#pragma unroll
        for (int v = 0; v < V; v++) {
            out[((size_t)(0 * V + v) * n + p) * stride + f0] = Dm[v];
            if (P::kNcp) out[((size_t)(1 * V + v) * n + p) * stride + f0] = Dp[v];
        }
    }
}

// First passes of get_subcell_data / get_lobatto_data for one cell held in LDS (uhL [v][NN])
template <class P, int N>
ADG_DEV void cell_ranges_stage1(const double *u2l, const double *u2b, const double *uhL, double *tyL, double *byL,
                                int tt, int nthr)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, V = S::V;
    for (int task = tt; task < n * s; task += nthr) {  // Euler/limiter.f90:25-27
        const int a = task % n, qq = task / n;
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int bb = 0; bb < n; bb++) acc = acc + uhL[v * NN + a + n * bb] * u2l[qq + s * bb];
            tyL[v * n * s + a + n * qq] = acc;
        }
    }
    for (int task = tt; task < NN; task += nthr) {  // :56-58
        const int a = task % n, qq = task / n;
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int bb = 0; bb < n; bb++) acc = acc + uhL[v * NN + a + n * bb] * u2b[qq + n * bb];
            byL[v * NN + a + n * qq] = acc;
        }
    }
}

// Second passes; returns true if a subcell average is inadmissible (Euler/limiter.f90:227-241)
template <class P, int N, class M>
ADG_DEV bool cell_ranges_stage2(const double *u2l, const double *u2b, const M &m, const double *tyL,
                                const double *byL, double *limL, double *lobL, int tt, int nthr)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, V = S::V;
    bool bad = false;
    for (int task = tt; task < SS; task += nthr) {  // :33-35
        const int p = task % s, qq = task / s;
        double L[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < n; a++) acc = acc + tyL[v * n * s + a + n * qq] * u2l[p + s * a];
            limL[v * SS + task] = acc;
            L[v] = acc;
        }
        bad = bad || P::inadmissible(m, L);
    }
    for (int task = tt; task < NN; task += nthr) {  // :64-66
        const int p = task % n, qq = task / n;
#pragma unroll
        for (int v = 0; v < V; v++) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < n; a++) acc = acc + byL[v * NN + a + n * qq] * u2b[p + n * a];
            lobL[v * NN + task] = acc;
        }
    }
    return bad;
}

// Pool slot of the subcell averages of a cell (StepBufs::lh), or -1 for a cell without
// limiter (and for a limited cell of a solve whose pool segment ran out: that solve has failed).
ADG_DEV int lim_slot(const StepBufs &b, int par, size_t cell)
{
    return b.status[par][cell] != 0 ? b.lh_slot[par][cell] : -1;
}

// DMP range test of one variable, Euler/limiter.f90:198-223 (argmax = 0): true if violated
ADG_DEV bool dmp_violated(const double r[6], double lmn, double lmx)
{
    const double mn = f_min(f_min(r[0], r[4]), r[2]);
    const double mx = f_max(f_max(r[1], r[5]), r[3]);
    const double ldiff = f_max(f_max(1.0e-4, 1.0e-3 * (lmx - lmn)), 0.0);
    return mn < lmn - ldiff || mx > lmx + ldiff;
}

// time-step candidate of one node, Euler/calc_time_step.f90:29-42
template <class P, class M>
ADG_DEV double node_dt(const M &m, const double *Q)
{
    double denom = 0.0;
    denom = denom + P::maxabs_eig(m, Q, 1.0, 0.0) / m.dx;
    denom = denom + P::maxabs_eig(m, Q, 0.0, 1.0) / m.dy;
    return m.cflpnpm / denom;
}

// ---------------------------------------------------------------------------
// Shared-map variant of the update (the same results as k_update): the first
// passes of the subcell and Lobatto maps (limy, loby) are formed once per cell
// in LDS by the cell's threads, the per-thread (min, max) partials reuse that
// LDS once the maps are read, and the cell's threads reduce them.  More LDS per
// cell, fewer FMAs than k_update: faster where N is large relative to nVar.
// ---------------------------------------------------------------------------
template <class P, int N>
struct SharedUpdShape {
    using S = Shape<P, N>;
    static constexpr int NSRC = 4 * S::V + 1;  // nodal, Lobatto, subcell min, subcell max per v, dt
    static constexpr int MAPS = S::V * S::NN + S::V * S::n * S::s + S::V * S::NN;
    static constexpr int PART = MAPS > NSRC * S::NN ? MAPS : NSRC * S::NN;
    static constexpr int LDS = PART + S::NPART;
};

template <class P, int N, bool SEED>
__global__ __launch_bounds__(kUpdBlock<P>) void k_update_shared(StepBufs b)
{
    using S = Shape<P, N>;
    using U = SharedUpdShape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, CPW = S::CPWU, V = S::V, NP = S::NPART;
    __shared__ double lds[CPW][U::LDS];
    __shared__ int sflag[CPW];
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const bool own = lc < CPW;
    const int lcs = own ? lc : CPW - 1;
    const int ntab = ld_const(b.meta + kTabUpd);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
    const int2 w = ld_const(b.tab[kTabUpd] + e);
    const SolveState st = ld_const(b.state + w.x);
    if (SEED ? !st.seed : !st.active) continue;
    const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + w.x);
    const auto &m = sa.m;
    const int par = st.parity;
    const int cell = w.y + lc;                 // local cell of the solve
    const size_t g = (size_t)sa.cell0 + cell;  // cell of the group arrays
    const bool act = own && cell < m.ncell;
    const double *uh_cur = b.uh[par];
    double *suh = &lds[lcs][0];
    double *sty = suh + V * NN;     // limy [v][a + n q]
    double *sby = sty + V * n * s;  // loby
    double *spart = suh;            // [source][tt], after the maps are read
    double *sred = suh + U::PART;   // [slot]
    if (own && tt == 0) sflag[lc] = 0;
    // the DMP bounds of this step (lmin/lmax of get_min_max), loaded now so their
    // latency overlaps the update instead of following the range reduction
    constexpr int KV = (V + NN - 1) / NN;  // variables per thread in the DMP pass
    double pmn[KV], pmx[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = tt + k * NN;
        const bool ok = !SEED && act && v < V;
        pmn[k] = ok ? b.lmin[(size_t)v * b.G + g] : 0.0;
        pmx[k] = ok ? b.lmax[(size_t)v * b.G + g] : 0.0;
    }

    double un[V];
    if (SEED || !act) {
#pragma unroll
        for (int v = 0; v < V; v++) un[v] = act ? uh_cur[g * V * NN + V * tt + v] : (v == 0 || v == 3 ? 1.0 : 0.0);
    } else {
        const int ci = cell % m.imax, cj = cell / m.imax;
        const size_t nfx = b.GX, nfy = b.GY;
        const size_t fl = sa.fx0 + ci + (size_t)(m.imax + 1) * cj, fb = sa.fy0 + ci + (size_t)m.imax * cj;
        const int sp = P::kNcp ? 1 : 0;  // right/top faces take D+ (Baer-Nunziato/main.f90:68)
        const double wi = op->w[i], wj = op->w[j], dt = st.dt;
#pragma unroll
        for (int v = 0; v < V; v++) {
            double d = b.lduh[g * V * NN + V * tt + v];
            const double uo = uh_cur[g * V * NN + V * tt + v];
            const double Fl = b.xflux[((size_t)(0 * V + v) * n + j) * nfx + fl];
            const double Fr = b.xflux[((size_t)(sp * V + v) * n + j) * nfx + fl + 1];
            const double Gb = b.yflux[((size_t)(0 * V + v) * n + i) * nfy + fb];
            const double Gt = b.yflux[((size_t)(sp * V + v) * n + i) * nfy + fb + m.imax];
            // surface integral, Euler/surface_integral.f90:182-196
            d = d - wj / m.dx * (Fr * op->frc[i] - Fl * op->flc[i]);
            d = d - wi / m.dy * (Gt * op->frc[j] - Gb * op->flc[j]);
            d = d / (wi * wj);  // element update, Euler/element_update.f90:218-226
            un[v] = uo + dt * d;
        }
        // stored once every value is formed: a store to uh[par ^ 1] between the loads would
        // hold the next variable's loads behind it (the compiler cannot rule out aliasing)
#pragma unroll
        for (int v = 0; v < V; v++) b.uh[par ^ 1][g * V * NN + V * tt + v] = un[v];
    }
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) suh[v * NN + tt] = un[v];
    }
    const double dtc = act ? node_dt<P>(m, un) : __builtin_nan("");
    __syncthreads();
    if (own) cell_ranges_stage1<P, N>(op->u2l, op->u2b, suh, sty, sby, tt, NN);
    __syncthreads();
    // second passes folded into per-thread (min, max) of Lobatto and subcell values
    double b0[V], l0[V], l1[V];
    bool bad = false;
#pragma unroll
    for (int v = 0; v < V; v++) {
        b0[v] = l0[v] = l1[v] = __builtin_nan("");
    }
    if (act) {
        for (int task = tt; task < SS; task += NN) {  // get_subcell_data :33-35, positivity :227-241
            const int p = task % s, qq = task / s;
            double L[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + sty[v * n * s + a + n * qq] * op->u2l[p + s * a];
                L[v] = acc;
                l0[v] = fold_min(l0[v], acc);
                l1[v] = fold_max(l1[v], acc);
            }
            bad = bad || P::inadmissible(m, L);
        }
        {  // get_lobatto_data :64-66, one node per thread
            const int p = tt % n, qq = tt / n;
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + sby[v * NN + a + n * qq] * op->u2b[p + n * a];
                b0[v] = acc;
            }
        }
    }
    __syncthreads();  // limy / loby are read: their space takes the partials
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            spart[(0 * V + v) * NN + tt] = un[v];
            spart[(1 * V + v) * NN + tt] = b0[v];  // = b1: one Lobatto node per thread
            spart[(2 * V + v) * NN + tt] = l0[v];
            spart[(3 * V + v) * NN + tt] = l1[v];
        }
        spart[4 * V * NN + tt] = dtc;
    }
    if (bad && !SEED) atomicOr(&sflag[lc], 1);
    __syncthreads();
    for (int slot = tt; own && slot < NP; slot += NN) {  // reduce the partials of the cell's threads
        // slot kind / V: 0 min, 1 max nodal; 2 min, 3 max Lobatto; 4 min, 5 max subcell; 6V: dt
        const int kind = slot / V, v = slot - kind * V;
        const bool is_max = slot < 6 * V && (kind & 1);
        const int src = slot < 6 * V ? (kind < 4 ? kind >> 1 : kind - 2) * V + v : 4 * V;
        double r = __builtin_nan("");
        for (int k = 0; k < NN; k++) r = is_max ? fold_max(r, spart[src * NN + k]) : fold_min(r, spart[src * NN + k]);
        sred[slot] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = tt + k * NN;
        if (!act || v >= V) continue;
        const double r[6] = {sred[0 * V + v], sred[1 * V + v], sred[2 * V + v],
                             sred[3 * V + v], sred[4 * V + v], sred[5 * V + v]};
        if (!SEED && dmp_violated(r, pmn[k], pmx[k])) atomicOr(&sflag[lc], 2);
        // next step's get_min_max step 1 for a cell without limiter, Euler/limiter.f90:123-132
        double mn = f_min(f_min(r[0], r[2]), r[4]), mx = f_max(f_max(r[1], r[3]), r[5]);
        const int ls = SEED ? lim_slot(b, par, g) : -1;
        if (ls >= 0) {  // limited cell: range of its subcell data, :135-139
            const double *L = b.lh[par] + (size_t)ls * V * SS;
            mn = __builtin_nan("");
            mx = mn;
            for (int kk = 0; kk < SS; kk++) {
                mn = fold_min(mn, L[v + V * kk]);
                mx = fold_max(mx, L[v + V * kk]);
            }
        }
        b.lmin_old[(size_t)v * b.G + g] = mn;
        b.lmax_old[(size_t)v * b.G + g] = mx;
    }
    if (act && tt == NN - 1) b.cell_dt[g] = sred[6 * V];
    __syncthreads();
    if (!SEED && act && tt == 0) {
        const int tr = sflag[lc] != 0;
        b.rc[g] = tr;
        if (tr) atomicAdd(&b.counters[kCntTroubled], 1);
    }
    __syncthreads();
    }
}

// Order-preserving 64-bit key of a double that is not NaN: unsigned key order is
// the numeric order, with -0 below +0, so an LDS atomic min/max of keys takes the
// min/max of the values independently of the order the threads arrive in.
ADG_DEV unsigned long long ord_key(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
// Inverse of ord_key; the untouched slot values (~0 for a min, 0 for a max) give NaN,
// the result of MINVAL/MAXVAL folds that met only NaN.
ADG_DEV double key_val(unsigned long long k)
{
    if (k == ~0ull || k == 0ull) return __builtin_nan("");
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}
// A NaN folds in as the slot's untouched value (~0 for a min, 0 for a max), which leaves
// the slot as it is: no branch around the atomic.
ADG_DEV void key_min(unsigned long long *slot, double x)
{
    atomicMin(slot, __builtin_isnan(x) ? ~0ull : ord_key(x));
}
ADG_DEV void key_max(unsigned long long *slot, double x)
{
    atomicMax(slot, __builtin_isnan(x) ? 0ull : ord_key(x));
}

// ---------------------------------------------------------------------------
// surface integral + element update + DMP (SEED = false), or only the
// per-cell limiter seeds of a freshly uploaded solution (SEED = true).
// Thread (cell, i, j) updates node (i,j).  The ranges the DMP needs (min/max
// of nodal, Lobatto and subcell values) come from the cell's nodal values in
// LDS: each thread forms its node's Lobatto value and, for a row (or half row)
// of the subcell grid, the row's first-pass values and then its points, in the
// order of get_subcell_data / get_lobatto_data, and folds its values into
// per-cell slots (four copies each) with LDS atomics on order-preserving keys.
// The operator rows the threads index live in LDS as well, and for
// Baer-Nunziato the entry's uh and lduh rows are staged there (16-byte row
// loads and stores instead of node-strided ones), and at N >= 3 its face lines
// (one lane per value instead of one load per node).
// ---------------------------------------------------------------------------
// waves per SIMD asked of k_update: Euler N=3 fits four without spilling (one more than
// the compiler's own choice, 1.22 -> 1.08 ms on 1024x1024); elsewhere the compiler's choice
template <class P, int N>
constexpr int kUpdWaves = (P::V == 4 && N == 3) ? 4 : (P::V == 4 ? 1 : 2);
// contiguous row staging of uh / lduh through LDS in k_update: Baer-Nunziato only (512x512:
// N=3 update 0.80 -> 0.73 ms, N=2 0.48 -> 0.42 ms; the Euler updates measured slower with it)
template <class P>
constexpr bool kUpdRows = P::V != 4;

template <class P, int N, bool SEED>
__global__ __launch_bounds__(kUpdBlock<P>) __attribute__((amdgpu_waves_per_eu(kUpdWaves<P, N>))) void k_update(StepBufs b)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, CPW = S::CPWU, V = S::V;
    constexpr int NP = 2 * V;  // LDS range slots per cell: subcell min (v), max (V + v)
    // KC copies of every slot: the cell's threads spread their atomics over the copies
    // (thread tt takes copy tt % KC), so at most ceil(NN / KC) lanes meet on one address
    constexpr int KC = kUpdKeyCopies;
    // ROWS (Baer-Nunziato): srow holds uh (then the new solution) of the entry's cells in
    // the layout of the group arrays ([cell][node][v]), so that global memory is read and
    // written in whole contiguous rows, and ssuh first the entry's lduh rows; otherwise
    // each thread reads and writes its node's values directly.  ssuh then holds the new
    // nodal values [cell][v][node] that the range maps read.
    constexpr bool ROWS = kUpdRows<P>;
    // FACES (ROWS, i.e. Baer-Nunziato, N <= 3: at most seven cells per workgroup): the entry's
    // face lines are staged in LDS, one lane per value, instead of each node loading its own
    // row's values (at N=2 it measured slower until the rows streamed non-temporally: then
    // 5.4 -> 5.1 ms on the ensemble's N=2 arena)
    constexpr bool FACES = ROWS && CPW <= 7;
    __shared__ __align__(16) double ssuh[CPW][V * NN];
    __shared__ __align__(16) double srow[ROWS ? CPW : 1][ROWS ? V * NN : 1];
    __shared__ unsigned long long skey[CPW][KC][NP];
    __shared__ int sflag[CPW];
    __shared__ double sdt[CPW][NN];  // time-step candidates of the nodes
    // Lobatto values [v][node] of the cells (Baer-Nunziato: in srow once its rows are stored)
    __shared__ double slob[ROWS ? 1 : CPW][ROWS ? 1 : V * NN];
    __shared__ double sface[FACES ? 4 * V * n * CPW : 1];  // the entry's face lines
    // the operator rows the threads index by their own node or point, read from LDS
    // instead of the constant tables (a vector memory access each)
    __shared__ double sopw[n], sopfl[n], sopfr[n], sopu2b[NN], sopu2l[s * n];
    const OpPtr op = (OpPtr)b.op;
    const int tid = threadIdx.x, lc = tid / NN, tt = tid - lc * NN;
    const int i = tt % n, j = tt / n;
    const bool own = lc < CPW;
    const int lcs = own ? lc : CPW - 1;
    for (int k = tid; k < s * n; k += kUpdBlock<P>) {
        sopu2l[k] = op->u2l[k];
        if (k < NN) sopu2b[k] = op->u2b[k];
        if (k < n) {
            sopw[k] = op->w[k];
            sopfl[k] = op->flc[k];
            sopfr[k] = op->frc[k];
        }
    }
    __syncthreads();
    const int ntab = ld_const(b.meta + kTabUpd);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
    const int2 w = ld_const(b.tab[kTabUpd] + e);
    const SolveState st = ld_const(b.state + w.x);
    if (SEED ? !st.seed : !st.active) continue;
    const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + w.x);
    const auto &m = sa.m;
    const int par = st.parity;
    const int cell = w.y + lc;                 // local cell of the solve
    const size_t g = (size_t)sa.cell0 + cell;  // cell of the group arrays
    const bool act = own && cell < m.ncell;
    const double *uh_cur = b.uh[par];
    double *suh = &ssuh[lcs][0];
    double *lobL;
    if constexpr (ROWS)
        lobL = &srow[lcs][0];
    else
        lobL = &slob[lcs][0];
    unsigned long long *key = &skey[lcs][tt % KC][0];  // [v] subcell min, [V + v] subcell max
    if (own) {
        if (tt == 0) sflag[lc] = 0;
        for (int k = tt; k < KC * NP; k += NN) {
            const int slot = k % NP;
            skey[lc][k / NP][slot] = slot >= V ? 0ull : ~0ull;
        }
    }
    // value of a slot: the min (max) over the copies of the order-preserving keys
    auto slot_val = [&](int slot) {
        const bool mx = slot >= V;
        unsigned long long r = skey[lcs][0][slot];
#pragma unroll
        for (int c = 1; c < KC; c++) {
            const unsigned long long k = skey[lcs][c][slot];
            r = mx ? (k > r ? k : r) : (k < r ? k : r);
        }
        return key_val(r);
    };
    // the DMP bounds of this step (lmin/lmax of get_min_max), loaded now so their
    // latency overlaps the update instead of following the range reduction
    constexpr int KV = (V + NN - 1) / NN;  // variables per thread in the DMP pass
    double pmn[KV], pmx[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = tt + k * NN;
        const bool ok = !SEED && act && v < V;
        pmn[k] = ok ? b.lmin[(size_t)v * b.G + g] : 0.0;
        pmx[k] = ok ? b.lmax[(size_t)v * b.G + g] : 0.0;
    }

    double un[V], dtc;
    if constexpr (ROWS) {
    const int nvalid = m.ncell - w.y < CPW ? m.ncell - w.y : CPW;  // cells of this entry
    const double wi = sopw[i], wj = sopw[j];
    // Every load of the entry is issued before the first of them is used, so that the entry
    // waits for one memory round trip, not one per loop trip: the uh / lduh rows and (FACES)
    // the face lines into registers, then into LDS.  PAIRS: a cell's row has an even number
    // of doubles, so the rows of every entry start 16-byte aligned and move as double pairs.
    // Otherwise the rows move as pairs from the first 16-byte boundary: h leading doubles
    // (0 or 1) before it, and one trailing double when the rest is odd.
    constexpr bool PAIRS = (V * NN) % 2 == 0;
    constexpr int KR = ((CPW * V * NN) / 2 + kUpdBlock<P> - 1) / kUpdBlock<P>;  // row pairs per lane
    const int cnt = nvalid * V * NN;
    const size_t base = ((size_t)sa.cell0 + w.y) * V * NN;
    const int h = PAIRS ? 0 : (int)(base & 1), np = (cnt - h) / 2;
    double2 ru[KR], rl[KR];
    {
        const double2 *u2 = (const double2 *)(uh_cur + base + h), *l2 = (const double2 *)(b.lduh + base + h);
#pragma unroll
        for (int q = 0; q < KR; q++) {
            const int k = min(tid + q * kUpdBlock<P>, np - 1);  // lanes past the end reload the last pair
            // the rows stream through: non-temporal, so that they do not evict the face lines
            // the entries of a row share in L2 (the N=3 Baer-Nunziato update read 1.7x its
            // algorithmic bytes with plain loads, 1.06x with these)
            ru[q] = nt_load2(u2 + k);
            if (!SEED) rl[q] = nt_load2(l2 + k);
        }
    }
    // the odd ends of unpaired rows: the leading double (h = 1) and the trailing one
    const int ee = tid == 0 ? 0 : cnt - 1;
    const bool end_lane = !PAIRS && ((tid == 0 && h) || (tid == 1 && ((cnt - h) & 1)));
    double eu = 0.0, el = 0.0;
    if (end_lane) {
        eu = uh_cur[base + ee];
        if (!SEED) el = b.lduh[base + ee];
    }
    // the surface-integral terms of this thread's node (one load per node and value),
    // Euler/surface_integral.f90:182-196
    double sx[V], sy[V];
    if (!FACES && !SEED && act) {
        const int ci = cell % m.imax, cj = cell / m.imax;
        const size_t nfx = b.GX, nfy = b.GY;
        const size_t fl = sa.fx0 + ci + (size_t)(m.imax + 1) * cj, fb = sa.fy0 + ci + (size_t)m.imax * cj;
        const int sp = P::kNcp ? 1 : 0;  // right/top faces take D+ (Baer-Nunziato/main.f90:68)
#pragma unroll
        for (int v = 0; v < V; v++) {
            const double Fl = b.xflux[((size_t)(0 * V + v) * n + j) * nfx + fl];
            const double Fr = b.xflux[((size_t)(sp * V + v) * n + j) * nfx + fl + 1];
            const double Gb = b.yflux[((size_t)(0 * V + v) * n + i) * nfy + fb];
            const double Gt = b.yflux[((size_t)(sp * V + v) * n + i) * nfy + fb + m.imax];
            sx[v] = wj / m.dx * (Fr * sopfr[i] - Fl * sopfl[i]);
            sy[v] = wi / m.dy * (Gt * sopfr[j] - Gb * sopfl[j]);
        }
    }
    // FACES: the face lines of the entry's cells in LDS, one lane per value: a cell's n
    // nodes of a row (column) read the same left/right (bottom/top) value, which the loads
    // above issue n times.  sface[side][v][p][cell]; sides left, right, bottom, top.
    // Lane tid takes cell c = tid % CPW of the entry and the values r = tid / CPW + q FG of
    // its lines (r = (side V + v) n + p), so its faces are located once.
    constexpr int FG = kUpdBlock<P> / CPW;                            // lanes per cell
    constexpr int FR = FACES ? (4 * V * n + FG - 1) / FG : 1;         // face values per lane
    const int fc = tid % CPW;
    const bool face_lane = FACES && !SEED && tid < FG * CPW && fc < nvalid;
    double xs[FR];
    if (face_lane) {
        const size_t nfx = b.GX, nfy = b.GY;
        const int sp = P::kNcp ? 1 : 0;  // right/top faces take D+ (Baer-Nunziato/main.f90:68)
        const int lcell = w.y + fc, ci = lcell % m.imax, cj = lcell / m.imax;
        const size_t fl = sa.fx0 + ci + (size_t)(m.imax + 1) * cj, fb = sa.fy0 + ci + (size_t)m.imax * cj;
#pragma unroll
        for (int q = 0; q < FR; q++) {
            const int r = tid / CPW + q * FG;
            if (r < 4 * V * n) {
                const int p = r % n, fv = r / n, v = fv % V, f = fv / V;
                if (f < 2)
                    xs[q] = b.xflux[((size_t)((f ? sp : 0) * V + v) * n + p) * nfx + fl + f];
                else
                    xs[q] = b.yflux[((size_t)((f == 3 ? sp : 0) * V + v) * n + p) * nfy + fb + (f == 3 ? m.imax : 0)];
            }
        }
    }
    // now into LDS, unguarded so that the compiler cannot sink each load into a guarded
    // store: a lane past the end writes the last pair again, the value its owner writes
    double *flat = &srow[0][0], *sl = &ssuh[0][0];
#pragma unroll
    for (int q = 0; q < KR; q++) {
        const int k = min(tid + q * kUpdBlock<P>, np - 1);
        if constexpr (PAIRS) {
            ((double2 *)flat)[k] = ru[q];
            if (!SEED) ((double2 *)sl)[k] = rl[q];
        } else {
            flat[h + 2 * k] = ru[q].x;
            flat[h + 2 * k + 1] = ru[q].y;
            if (!SEED) {
                sl[h + 2 * k] = rl[q].x;
                sl[h + 2 * k + 1] = rl[q].y;
            }
        }
    }
    if (end_lane) {
        flat[ee] = eu;
        if (!SEED) sl[ee] = el;
    }
    // the face lines into LDS after the rows: the row stores wait for loads issued after the
    // face loads, so that both arrive in one round trip
    if (face_lane) {
#pragma unroll
        for (int q = 0; q < FR; q++) {
            const int r = tid / CPW + q * FG;
            if (r < 4 * V * n) sface[r * CPW + fc] = xs[q];
        }
    }
    __syncthreads();  // the entry's uh and lduh rows (and face lines) in LDS
    if (SEED || !act) {
#pragma unroll
        for (int v = 0; v < V; v++) un[v] = act ? srow[lcs][V * tt + v] : (v == 0 || v == 3 ? 1.0 : 0.0);
    } else {
        const double dt = st.dt;
        constexpr int FS = V * n * CPW;  // doubles per side in sface
#pragma unroll
        for (int v = 0; v < V; v++) {
            if constexpr (FACES) {  // surface-integral terms, Euler/surface_integral.f90:182-196
                const double Fl = sface[0 * FS + (v * n + j) * CPW + lc];
                const double Fr = sface[1 * FS + (v * n + j) * CPW + lc];
                const double Gb = sface[2 * FS + (v * n + i) * CPW + lc];
                const double Gt = sface[3 * FS + (v * n + i) * CPW + lc];
                sx[v] = wj / m.dx * (Fr * sopfr[i] - Fl * sopfl[i]);
                sy[v] = wi / m.dy * (Gt * sopfr[j] - Gb * sopfl[j]);
            }
            double d = ssuh[lc][V * tt + v];
            const double uo = srow[lc][V * tt + v];
            d = d - sx[v];
            d = d - sy[v];
            d = d / (wi * wj);  // element update, Euler/element_update.f90:218-226
            un[v] = uo + dt * d;
            srow[lc][V * tt + v] = un[v];  // this thread's own node, read just above
        }
    }
    __syncthreads();  // every thread has read its lduh row: ssuh takes the new values
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) suh[v * NN + tt] = un[v];
    }
    dtc = act ? node_dt<P>(m, un) : __builtin_nan("");
    if (own) sdt[lc][tt] = dtc;
    __syncthreads();  // new nodal values and slot initial values in place
    if (!SEED) {
        if constexpr (PAIRS) {
            double2 *o2 = (double2 *)(b.uh[par ^ 1] + base);
            const double2 *f2 = (const double2 *)flat;
#pragma unroll 2
            for (int k = tid; k < cnt / 2; k += kUpdBlock<P>) nt_store2(o2 + k, f2[k]);
        } else {
            double2 *o2 = (double2 *)(b.uh[par ^ 1] + base + h);
#pragma unroll 2
            for (int k = tid; k < np; k += kUpdBlock<P>) nt_store2(o2 + k, make_double2(flat[h + 2 * k], flat[h + 2 * k + 1]));
            const int e = tid == 0 ? 0 : cnt - 1;
            if ((tid == 0 && h) || (tid == 1 && ((cnt - h) & 1))) b.uh[par ^ 1][base + e] = flat[e];
        }
    }
    __syncthreads();  // the rows are read: srow takes the Lobatto values
    } else {
    if (SEED || !act) {
#pragma unroll
        for (int v = 0; v < V; v++) un[v] = act ? uh_cur[g * V * NN + V * tt + v] : (v == 0 || v == 3 ? 1.0 : 0.0);
    } else {
        const int ci = cell % m.imax, cj = cell / m.imax;
        const size_t nfx = b.GX, nfy = b.GY;
        const size_t fl = sa.fx0 + ci + (size_t)(m.imax + 1) * cj, fb = sa.fy0 + ci + (size_t)m.imax * cj;
        const int sp = P::kNcp ? 1 : 0;  // right/top faces take D+ (Baer-Nunziato/main.f90:68)
        const double wi = sopw[i], wj = sopw[j], dt = st.dt;
#pragma unroll
        for (int v = 0; v < V; v++) {
            double d = b.lduh[g * V * NN + V * tt + v];
            const double uo = uh_cur[g * V * NN + V * tt + v];
            const double Fl = b.xflux[((size_t)(0 * V + v) * n + j) * nfx + fl];
            const double Fr = b.xflux[((size_t)(sp * V + v) * n + j) * nfx + fl + 1];
            const double Gb = b.yflux[((size_t)(0 * V + v) * n + i) * nfy + fb];
            const double Gt = b.yflux[((size_t)(sp * V + v) * n + i) * nfy + fb + m.imax];
            // surface integral, Euler/surface_integral.f90:182-196
            d = d - wj / m.dx * (Fr * sopfr[i] - Fl * sopfl[i]);
            d = d - wi / m.dy * (Gt * sopfr[j] - Gb * sopfl[j]);
            d = d / (wi * wj);  // element update, Euler/element_update.f90:218-226
            un[v] = uo + dt * d;
        }
        // stored once every value is formed: a store to uh[par ^ 1] between the loads would
        // hold the next variable's loads behind it (the compiler cannot rule out aliasing)
#pragma unroll
        for (int v = 0; v < V; v++) b.uh[par ^ 1][g * V * NN + V * tt + v] = un[v];
    }
    if (own) {
#pragma unroll
        for (int v = 0; v < V; v++) suh[v * NN + tt] = un[v];
    }
    dtc = act ? node_dt<P>(m, un) : __builtin_nan("");
    if (own) sdt[lc][tt] = dtc;
    __syncthreads();  // nodal values and slot initial values in place
    }
    bool bad = false;
    if (act) {
        // (the nodal ranges and the time-step minimum are folded from LDS below)
        {  // get_lobatto_data :56-66 at this thread's node (p, q) = (i, j)
            double ub_q[n], ub_p[n];
#pragma unroll
            for (int k = 0; k < n; k++) {
                ub_q[k] = sopu2b[j + n * k];
                ub_p[k] = sopu2b[i + n * k];
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) {
                    double ly = 0.0;  // loby(a, q), :56-58
#pragma unroll
                    for (int bb = 0; bb < n; bb++) ly = ly + suh[v * NN + a + n * bb] * ub_q[bb];
                    acc = acc + ly * ub_p[a];  // :64-66
                }
                lobL[v * NN + tt] = acc;
                __builtin_amdgcn_sched_barrier(0);  // one variable's loads in flight at a time
            }
        }
        // get_subcell_data :25-35 and positivity :227-241 by row pieces: thread tt < NC takes
        // row q of the subcell grid (or half of it, when 2s rows fit the cell's threads), forms
        // the row's first-pass values limy(:, q) once for every variable and then the subcell
        // values of its points from them (the same sums, in the same order)
        double l0[V], l1[V];
#pragma unroll
        for (int v = 0; v < V; v++) l0[v] = l1[v] = __builtin_nan("");
        constexpr bool HALF = 2 * s <= NN;
        constexpr int NC = HALF ? 2 * s : s;
        if (tt < NC) {
            const int qq = HALF ? tt % s : tt, h = HALF ? tt / s : 0;
            const int pa = HALF ? h * n : 0, pb = HALF ? (h ? s : n) : s;
            double cq[n];
#pragma unroll
            for (int k = 0; k < n; k++) cq[k] = sopu2l[qq + s * k];
            double ly[V][n];  // limy(a, q), :25-27
#pragma unroll
            for (int v = 0; v < V; v++) {
#pragma unroll
                for (int a = 0; a < n; a++) ly[v][a] = 0.0;
#pragma unroll
                for (int bb = 0; bb < n; bb++)
#pragma unroll
                    for (int a = 0; a < n; a++) ly[v][a] = ly[v][a] + suh[v * NN + a + n * bb] * cq[bb];
            }
#pragma unroll 1
            for (int p = pa; p < pb; p++) {
                double cp[n], L[V];
#pragma unroll
                for (int k = 0; k < n; k++) cp[k] = sopu2l[p + s * k];
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double acc = 0.0;  // :33-35
#pragma unroll
                    for (int a = 0; a < n; a++) acc = acc + ly[v][a] * cp[a];
                    L[v] = acc;
                    l0[v] = fold_min(l0[v], acc);
                    l1[v] = fold_max(l1[v], acc);
                }
                bad = bad || P::inadmissible(m, L);
            }
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            key_min(&key[v], l0[v]);
            key_max(&key[V + v], l1[v]);
        }
    }
    if (bad && !SEED) atomicOr(&sflag[lc], 1);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = tt + k * NN;
        if (!act || v >= V) continue;
        // nodal range, Euler/limiter.f90:123-132, from the new nodal values in LDS
        double u0 = __builtin_nan(""), u1 = u0;
#pragma unroll
        for (int k2 = 0; k2 < NN; k2++) {
            u0 = fold_min(u0, suh[v * NN + k2]);
            u1 = fold_max(u1, suh[v * NN + k2]);
        }
        double b0 = __builtin_nan(""), b1 = b0;  // Lobatto range, :64-66
#pragma unroll
        for (int k2 = 0; k2 < NN; k2++) {
            b0 = fold_min(b0, lobL[v * NN + k2]);
            b1 = fold_max(b1, lobL[v * NN + k2]);
        }
        const double r[6] = {u0, u1, b0, b1, slot_val(v), slot_val(V + v)};
        if (!SEED && dmp_violated(r, pmn[k], pmx[k])) atomicOr(&sflag[lc], 2);
        // next step's get_min_max step 1 for a cell without limiter, Euler/limiter.f90:123-132
        double mn = f_min(f_min(r[0], r[2]), r[4]), mx = f_max(f_max(r[1], r[3]), r[5]);
        const int ls = SEED ? lim_slot(b, par, g) : -1;
        if (ls >= 0) {  // limited cell: range of its subcell data, :135-139
            const double *L = b.lh[par] + (size_t)ls * V * SS;
            mn = __builtin_nan("");
            mx = mn;
            for (int kk = 0; kk < SS; kk++) {
                mn = fold_min(mn, L[v + V * kk]);
                mx = fold_max(mx, L[v + V * kk]);
            }
        }
        b.lmin_old[(size_t)v * b.G + g] = mn;
        b.lmax_old[(size_t)v * b.G + g] = mx;
    }
    if (act && tt == NN - 1) {  // the cell's time-step candidate, Euler/calc_time_step.f90:29-42
        double d = __builtin_nan("");
#pragma unroll
        for (int k2 = 0; k2 < NN; k2++) d = fold_min(d, sdt[lc][k2]);
        b.cell_dt[g] = d;
    }
    __syncthreads();
    if (!SEED && act && tt == 0) {
        const int tr = sflag[lc] != 0;
        b.rc[g] = tr;
        if (tr) atomicAdd(&b.counters[kCntTroubled], 1);
    }
    __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// spread_recompute + allocate_limiter + compaction (one thread per cell)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_spread(StepBufs b)
{
    __shared__ int wcount[kBlock / 64], wbase[kBlock / 64], bbase;
    const int ntab = ld_const(b.meta + kTabCell);
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabCell] + e);
        const SolveState st = ld_const(b.state + w.x);
        if (!st.active) continue;
        const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + w.x);
        const auto &m = sa.m;
        const int par = st.parity;
        const int cell = w.y + threadIdx.x;  // local cell of the solve
        const int g = sa.cell0 + cell;
        int flag = 0;
        if (cell < m.ncell) {
            if (b.rc[g] == 1) {
                flag = 1;
            } else {
                // a face neighbour marked 1 marks this cell 2 (Euler/limiter.f90:336-350; the
                // face-neighbour relation of the clamped/wrapped table is symmetric)
                const int ci = cell % m.imax, cj = cell / m.imax;
                const int nbs[4] = {nb_index(ci, -1, m.imax, m.wrap & 1) + m.imax * cj,
                                    nb_index(ci, 1, m.imax, m.wrap & 1) + m.imax * cj,
                                    ci + m.imax * nb_index(cj, -1, m.jmax, m.wrap >> 1),
                                    ci + m.imax * nb_index(cj, 1, m.jmax, m.wrap >> 1)};
#pragma unroll
                for (int f = 0; f < 4; f++)
                    if (nbs[f] != cell && b.rc[sa.cell0 + nbs[f]] == 1) flag = 2;
            }
            b.rc_final[g] = flag;
            b.status[par ^ 1][g] = flag == 1 ? 1 : 0;  // allocate_limiter, Euler/limiter.f90:308-320
        }
        const bool take = flag > 0;
        const unsigned long long mask = __ballot(take);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int below = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[wave] = __popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int k = 0; k < kBlock / 64; k++) {
                wbase[k] = tot;
                tot += wcount[k];
            }
            bbase = tot ? atomicAdd(&b.counters[kCntList], tot) : 0;
        }
        __syncthreads();
        if (take) b.list[bbase + wbase[wave] + below] = g;
        __syncthreads();
    }
}

// The operator tables the subcell passes index per lane (u2l: subcell averages from nodal
// values, l2u: nodal values from subcell averages, u2b: Lobatto values), copied to LDS once per
// workgroup: read from global memory, each of the dozen barrier-separated passes of a cell
// waited for them (Baer-Nunziato N=3 subcell recompute: a third of a cell's time)
struct SubOps {
    const double *u2l, *l2u, *u2b;
};
template <int N>
struct SubOpLds {
    static constexpr int n = N + 1, s = 2 * N + 1;
    double u2l[s * n], l2u[n * s], u2b[n * n];
};
template <int N>
ADG_DEV SubOps stage_subops(OpPtr op, SubOpLds<N> &d, int lane, int nthr)
{
    constexpr int n = N + 1, s = 2 * N + 1;
    for (int k = lane; k < s * n; k += nthr) {
        d.u2l[k] = op->u2l[k];
        d.l2u[k] = op->l2u[k];
    }
    for (int k = lane; k < n * n; k += nthr) d.u2b[k] = op->u2b[k];
    __syncthreads();
    return SubOps{d.u2l, d.l2u, d.u2b};
}

// Common tail of both subcell kernels: put_subcell_data (Euler/limiter.f90:86-96),
// re-DMP of cells marked 2 (:619-638), next-step limiter seeds and dt candidate.
// NT: threads of the workgroup (one cell per workgroup); sNth: LDS word that broadcasts the
// pool slot count of a limited cell to the workgroup
template <class P, int N, int NT, class M>
ADG_DEV void subcell_finish(const M &m, const StepBufs &b, size_t cell, int sol, int lh0, int lhcap, int par, int flag,
                            const double *sSub, double *sUh, double *sLim, double *sLob, double *sDt, double *sRng,
                            int *sFlag, int *sNth, const SubOps &o)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, V = S::V;
    const int lane = threadIdx.x;
    // a cell marked 1 is limited whatever the re-DMP finds: its pool slot is taken now, so
    // that the atomic's round trip overlaps the passes below instead of ending the cell
    int nth1 = 0;
    if (flag == 1 && lane == 0) nth1 = atomicAdd(b.lh_count + sol, 1);
    // this step's DMP bounds of the cell, loaded now so their latency overlaps the maps
    constexpr int KV = (V + NT - 1) / NT;
    double pmn[KV], pmx[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = lane + k * NT;
        pmn[k] = v < V ? b.lmin[(size_t)v * b.G + cell] : 0.0;
        pmx[k] = v < V ? b.lmax[(size_t)v * b.G + cell] : 0.0;
    }
    // Each pair of separable passes below runs as one pass per row: a thread forms the row's
    // first-pass sums once and then the row's outputs from them (the same sums in the same
    // order, so the same bits), one barrier and one LDS round trip fewer per pair.
    // put_subcell_data (Euler/limiter.f90:86-96): the nodal values of the new subcell
    // averages, node row bb of variable v per thread
    for (int task = lane; task < V * n; task += NT) {
        const int v = task / n, bb = task - v * n;
        double y[s];  // :86-90, the first pass at (p, bb)
#pragma unroll
        for (int p = 0; p < s; p++) {
            double acc = 0.0;
#pragma unroll
            for (int qq = 0; qq < s; qq++) acc = acc + sSub[v * SS + p + s * qq] * o.l2u[bb + n * qq];
            y[p] = acc;
        }
#pragma unroll
        for (int a = 0; a < n; a++) {
            double acc = 0.0;
#pragma unroll
            for (int p = 0; p < s; p++) acc = acc + y[p] * o.l2u[a + n * p];
            sUh[v * NN + a + n * bb] = acc;
            b.uh[par ^ 1][cell * V * NN + V * (a + n * bb) + v] = acc;
        }
    }
    if (lane == 0) *sFlag = 0;
    __syncthreads();
    // the new cell's subcell data (get_subcell_data :25-35) and Lobatto data
    // (get_lobatto_data :56-66), a row of one variable per thread, and its nodes' time-step
    // candidates
    for (int task = lane; task < V * s + V * n + NN; task += NT) {
        if (task < V * s) {
            const int v = task / s, qq = task - v * s;
            double ty[n];  // limy(a, qq), :25-27
#pragma unroll
            for (int a = 0; a < n; a++) {
                double acc = 0.0;
#pragma unroll
                for (int bb = 0; bb < n; bb++) acc = acc + sUh[v * NN + a + n * bb] * o.u2l[qq + s * bb];
                ty[a] = acc;
            }
#pragma unroll
            for (int p = 0; p < s; p++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + ty[a] * o.u2l[p + s * a];
                sLim[v * SS + p + s * qq] = acc;
            }
        } else if (task < V * s + V * n) {
            const int t = task - V * s, v = t / n, qq = t - v * n;
            double by[n];  // loby(a, qq), :56-58
#pragma unroll
            for (int a = 0; a < n; a++) {
                double acc = 0.0;
#pragma unroll
                for (int bb = 0; bb < n; bb++) acc = acc + sUh[v * NN + a + n * bb] * o.u2b[qq + n * bb];
                by[a] = acc;
            }
#pragma unroll
            for (int p = 0; p < n; p++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + by[a] * o.u2b[p + n * a];
                sLob[v * NN + p + n * qq] = acc;
            }
        } else {
            const int node = task - V * s - V * n;
            double un[V];
#pragma unroll
            for (int v = 0; v < V; v++) un[v] = sUh[v * NN + node];
            sDt[node] = node_dt<P>(m, un);
        }
    }
    __syncthreads();
    // every range the rest needs, each fold over its group in element order (MINVAL / MAXVAL,
    // Euler/limiter.f90:123-132): sRng[kind * V + v],
    // kind 0/1 nodal min/max, 2/3 Lobatto, 4/5 subcell, 6/7 the new subcell averages (a limited
    // cell's next ranges, :135-139); sRng[8V] the time-step minimum.  A thread takes the min and
    // the max of one group (two independent chains), its elements loaded ahead of the folds
    // (compile-time counts, fully unrolled).  With two waves the first folds and the second
    // checks the positivity of the new subcell averages (:227-241) point by point.
    bool bad = false;
    auto fold_task = [&](int t) {  // t < 4V + 1: subcell groups, nodal / Lobatto groups, dt
        if (t < 2 * V) {
            const int grp = t / V, v = t - grp * V;
            const double *src = grp == 0 ? sLim + v * SS : sSub + v * SS;
            double r0 = __builtin_nan(""), r1 = r0;
            constexpr int C = SS < 25 ? SS : 25;
#pragma unroll
            for (int k0 = 0; k0 < SS; k0 += C) {
                double x[C];
#pragma unroll
                for (int k = 0; k < C; k++) x[k] = k0 + k < SS ? src[k0 + k] : 0.0;
#pragma unroll
                for (int k = 0; k < C && k0 + k < SS; k++) {
                    r0 = fold_min(r0, x[k]);
                    r1 = fold_max(r1, x[k]);
                }
            }
            sRng[(4 + 2 * grp) * V + v] = r0;
            sRng[(5 + 2 * grp) * V + v] = r1;
        } else if (t < 4 * V) {
            const int u = t - 2 * V, grp = u / V, v = u - grp * V;
            const double *src = grp == 0 ? sUh + v * NN : sLob + v * NN;
            double x[NN], r0 = __builtin_nan(""), r1 = r0;
#pragma unroll
            for (int k = 0; k < NN; k++) x[k] = src[k];
#pragma unroll
            for (int k = 0; k < NN; k++) {
                r0 = fold_min(r0, x[k]);
                r1 = fold_max(r1, x[k]);
            }
            sRng[(2 * grp) * V + v] = r0;
            sRng[(2 * grp + 1) * V + v] = r1;
        } else {
            double r = __builtin_nan("");
#pragma unroll
            for (int k = 0; k < NN; k++) r = fold_min(r, sDt[k]);
            sRng[8 * V] = r;
        }
    };
    auto positivity_task = [&](int pt) {  // subcell point pt
        double L[V];
#pragma unroll
        for (int v = 0; v < V; v++) L[v] = sLim[v * SS + pt];
        bad = bad || P::inadmissible(m, L);
    };
    if constexpr (NT >= 128) {
        if (lane < 64) {
            for (int t = lane; t < 4 * V + 1; t += 64) fold_task(t);
        } else {
            for (int pt = lane - 64; pt < SS; pt += NT - 64) positivity_task(pt);
        }
    } else {
        for (int t = lane; t < 4 * V + 1 + SS; t += NT) {
            if (t < 4 * V + 1)
                fold_task(t);
            else
                positivity_task(t - 4 * V - 1);
        }
    }
    if (bad) atomicOr(sFlag, 1);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int v = lane + k * NT;
        if (v >= V) continue;
        const double r[6] = {sRng[v], sRng[V + v], sRng[2 * V + v], sRng[3 * V + v], sRng[4 * V + v], sRng[5 * V + v]};
        if (dmp_violated(r, pmn[k], pmx[k])) atomicOr(sFlag, 2);
    }
    __syncthreads();
    const int status = (flag == 1) ? 1 : (*sFlag != 0 ? 1 : 0);
    for (int v = lane; v < V; v += NT) {
        double mn, mx;
        if (status) {  // next get_min_max uses Lh (:135-139)
            mn = sRng[6 * V + v];
            mx = sRng[7 * V + v];
        } else {
            mn = f_min(f_min(sRng[v], sRng[2 * V + v]), sRng[4 * V + v]);
            mx = f_max(f_max(sRng[V + v], sRng[3 * V + v]), sRng[5 * V + v]);
        }
        b.lmin_old[(size_t)v * b.G + cell] = mn;
        b.lmax_old[(size_t)v * b.G + cell] = mx;
    }
    if (lane == NT - 1) {
        b.cell_dt[cell] = sRng[8 * V];
        b.status[par ^ 1][cell] = status;
    }
    if (status) {  // NewLh in a pool slot of the solve, and Lh = NewLh at update_limiter (:287)
        if (lane == 0) *sNth = flag == 1 ? nth1 : atomicAdd(b.lh_count + sol, 1);
        __syncthreads();  // status is the same in every thread of the workgroup
        const int nth = *sNth;
        const int ls = nth < lhcap ? lh0 + nth : -1;  // none left: the solve ends failed after this step
        if (lane == 0) b.lh_slot[par ^ 1][cell] = ls;
        if (ls >= 0) {
            double *L = b.lh[par ^ 1] + (size_t)ls * V * SS;
            for (int k = lane; k < V * SS; k += NT) L[k] = sSub[(k % V) * SS + k / V];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Euler subcell_recompute over the compacted list, one wavefront per cell
// (Euler/limiter.f90:411-647)
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(64) void k_subcell(StepBufs b)
{
    using P = EulerPDE;
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, V = S::V, H = s + 2, HH = H * H;
    __shared__ double sLy[5][V * n * s];
    __shared__ double sS0[V * HH];
    __shared__ double sSl[2][V * SS];
    __shared__ double sW[4][V * HH];
    __shared__ double sFx[V * (s + 1) * s];
    __shared__ double sFy[V * s * (s + 1)];
    __shared__ double sSub[V * SS];
    __shared__ double sUh[V * NN];
    __shared__ double sLim[V * SS], sLob[V * NN];
    __shared__ double sDt[NN];
    __shared__ double sRng[8 * V + 1];
    __shared__ int sFlag, sNth;
    const OpPtr op = (OpPtr)b.op;
    const int lane = threadIdx.x;
    __shared__ SubOpLds<N> sOps;
    const SubOps o = stage_subops<N>(op, sOps, lane, 64);
    const int count = b.counters[kCntList];
    for (int item = blockIdx.x; item < count; item += gridDim.x) {
        const int gcell = b.list[item];
        const int slot = ld_const(b.cell_solve + gcell);
        const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + slot);
        const auto &m = sa.m;
        const SolveState st = ld_const(b.state + slot);
        const int par = st.parity;
        const double dt = st.dt, time = st.time;
        const double ldx = m.dx / (double)s, ldy = m.dy / (double)s;
        const int cell = gcell - sa.cell0;  // local cell of the solve
        const int flag = b.rc_final[gcell];
        const int ci = cell % m.imax, cj = cell / m.imax;
        int nbk[5];  // cells of the group arrays
        nbk[0] = gcell;
        nbk[1] = sa.cell0 + nb_index(ci, -1, m.imax, m.wrap & 1) + m.imax * cj;
        nbk[2] = sa.cell0 + nb_index(ci, 1, m.imax, m.wrap & 1) + m.imax * cj;
        nbk[3] = sa.cell0 + ci + m.imax * nb_index(cj, -1, m.jmax, m.wrap >> 1);
        nbk[4] = sa.cell0 + ci + m.imax * nb_index(cj, 1, m.jmax, m.wrap >> 1);
        int stk[5];
#pragma unroll
        for (int k = 0; k < 5; k++) stk[k] = lim_slot(b, par, nbk[k]);  // pool slot, -1: no limiter
        // subcell data of olduh for unlimited blocks, first pass (Euler/limiter.f90:25-27, :476-477)
        for (int task = lane; task < 5 * n * s; task += 64) {
            const int blk = task / (n * s), r = task - blk * n * s, a = r % n, qq = r / n;
            if (stk[blk] >= 0) continue;
            const double *u = b.uh[par] + (size_t)nbk[blk] * V * NN;
#pragma unroll
            for (int v = 0; v < V; v++) {
                double acc = 0.0;
#pragma unroll
                for (int bb = 0; bb < n; bb++) acc = acc + u[v + V * (a + n * bb)] * o.u2l[qq + s * bb];
                sLy[blk][v * n * s + a + n * qq] = acc;
            }
        }
        __syncthreads();
        // subuh0 on the subgrid 0..s+1 (index P = ii): the cell and one ring of its face
        // neighbours' subcells (:470-485), with the DMR ghost overrides (:491-522)
        for (int task = lane; task < HH; task += 64) {
            const int ii = task % H, jj = task / H;
            const bool in_x = ii >= 1 && ii <= s, in_y = jj >= 1 && jj <= s;
            int blk, p, qq;
            if (in_x && in_y) { blk = 0; p = ii - 1; qq = jj - 1; }
            else if (ii == 0 && in_y) { blk = 1; p = s - 1; qq = jj - 1; }
            else if (ii == s + 1 && in_y) { blk = 2; p = 0; qq = jj - 1; }
            else if (jj == 0 && in_x) { blk = 3; p = ii - 1; qq = s - 1; }
            else if (jj == s + 1 && in_x) { blk = 4; p = ii - 1; qq = 0; }
            else {  // corner subcells are never read
#pragma unroll
                for (int v = 0; v < V; v++) sS0[v * HH + task] = 0.0;
                continue;
            }
            double val[V];
            if (stk[blk] < 0) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < n; a++) acc = acc + sLy[blk][v * n * s + a + n * qq] * o.u2l[p + s * a];
                    val[v] = acc;
                }
            } else {
                const double *L = b.lh[par] + (size_t)stk[blk] * V * SS;
#pragma unroll
                for (int v = 0; v < V; v++) val[v] = L[v + V * (p + s * qq)];
            }
            if (m.bc_mode == kBcDmr) {
                if (ci == 0 && ii == 0)
#pragma unroll
                    for (int v = 0; v < V; v++) val[v] = m.qa[v];
                if (cj == 0 && jj == 0 && (double)(ci + 1) * m.dx > 1.0 / 6.0) val[2] = -val[2];
                if (cj == m.jmax - 1 && jj == s + 1) {
                    const bool pre = (double)(ci + 1) * m.dx <= dmr_shock_x(m, time);
#pragma unroll
                    for (int v = 0; v < V; v++) val[v] = pre ? m.qa[v] : m.qb[v];
                }
            } else {  // the side codes (inflow / reflective) on ghosts outside the domain
                const int gx = (ci == 0 && ii == 0) ? -1 : ((ci == m.imax - 1 && ii == s + 1) ? 1 : 0);
                const int gy = (cj == 0 && jj == 0) ? -1 : ((cj == m.jmax - 1 && jj == s + 1) ? 1 : 0);
                side_subcell<P>(m, gx, gy, val);
            }
#pragma unroll
            for (int v = 0; v < V; v++) sS0[v * HH + task] = val[v];
        }
        __syncthreads();
#define S0(v, ii, jj) sS0[(v) * HH + (ii) + H * (jj)]
        // minmod slopes on the cell's own subcells, :529-536
        for (int task = lane; task < SS; task += 64) {
            const int ii = task % s + 1, jj = task / s + 1;
#pragma unroll
            for (int v = 0; v < V; v++) {
                sSl[0][v * SS + task] = minmod(S0(v, ii + 1, jj) - S0(v, ii, jj), S0(v, ii, jj) - S0(v, ii - 1, jj));
                sSl[1][v * SS + task] = minmod(S0(v, ii, jj + 1) - S0(v, ii, jj), S0(v, ii, jj) - S0(v, ii, jj - 1));
            }
        }
        __syncthreads();
        // Hancock half step, :540-569 (halo slopes are zero)
        for (int task = lane; task < HH; task += 64) {
            const int ii = task % H, jj = task / H;
            const bool in_x = ii >= 1 && ii <= s, in_y = jj >= 1 && jj <= s;
            if (!in_x && !in_y) continue;
            const bool inner = in_x && in_y;
            const int sidx = (ii - 1) + s * (jj - 1);
            double wLx[V], wRx[V], wLy[V], wRy[V], FLx[2 * V], FRx[2 * V], FLy[2 * V], FRy[2 * V], Qt[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double sl = inner ? sSl[0][v * SS + sidx] : 0.0;
                wLx[v] = S0(v, ii, jj) - 0.5 * sl;
                wRx[v] = S0(v, ii, jj) + 0.5 * sl;
            }
            P::flux(m, wLx, FLx);
            P::flux(m, wRx, FRx);
#pragma unroll
            for (int v = 0; v < V; v++) Qt[v] = -(FRx[v] - FLx[v]) / ldx;
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double sl = inner ? sSl[1][v * SS + sidx] : 0.0;
                wLy[v] = S0(v, ii, jj) - 0.5 * sl;
                wRy[v] = S0(v, ii, jj) + 0.5 * sl;
            }
            P::flux(m, wLy, FLy);
            P::flux(m, wRy, FRy);
#pragma unroll
            for (int v = 0; v < V; v++) Qt[v] = Qt[v] - (FRy[V + v] - FLy[V + v]) / ldy;
#pragma unroll
            for (int v = 0; v < V; v++) {
                sW[0][v * HH + task] = wLx[v] + 0.5 * dt * Qt[v];
                sW[1][v * HH + task] = wRx[v] + 0.5 * dt * Qt[v];
                sW[2][v * HH + task] = wLy[v] + 0.5 * dt * Qt[v];
                sW[3][v * HH + task] = wRy[v] + 0.5 * dt * Qt[v];
            }
        }
        __syncthreads();
        // subcell Rusanov fluxes, :571-588 (RusanovFluxS, :361-378)
        for (int task = lane; task < (s + 1) * s + s * (s + 1); task += 64) {
            const bool xdir = task < (s + 1) * s;
            const int t2 = xdir ? task : task - (s + 1) * s;
            double QL[V], QR[V], FL[2 * V], FR[2 * V];
            int ii, jj;
            if (xdir) {
                ii = t2 % (s + 1) + 1;
                jj = t2 / (s + 1) + 1;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    QL[v] = sW[1][v * HH + (ii - 1) + H * jj];
                    QR[v] = sW[0][v * HH + ii + H * jj];
                }
            } else {
                ii = t2 % s + 1;
                jj = t2 / s + 1;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    QL[v] = sW[3][v * HH + ii + H * (jj - 1)];
                    QR[v] = sW[2][v * HH + ii + H * jj];
                }
            }
            const double nx = xdir ? 1.0 : 0.0, ny = xdir ? 0.0 : 1.0;
            const double smax = f_max(P::maxabs_eig(m, QL, nx, ny), P::maxabs_eig(m, QR, nx, ny));
            P::flux(m, QL, FL);
            P::flux(m, QR, FR);
#pragma unroll
            for (int v = 0; v < V; v++) {
                double mm = 0.0;
                mm = mm + (FL[v] + FR[v]) * nx;
                mm = mm + (FL[V + v] + FR[V + v]) * ny;
                const double out = 0.5 * mm - 0.5 * smax * (QR[v] - QL[v]);
                if (xdir)
                    sFx[v * (s + 1) * s + t2] = out;
                else
                    sFy[v * s * (s + 1) + t2] = out;
            }
        }
        __syncthreads();
        // finite-volume update, :592-597
        for (int task = lane; task < SS; task += 64) {
            const int ii = task % s + 1, jj = task / s + 1;
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double *fx = sFx + v * (s + 1) * s, *fy = sFy + v * s * (s + 1);
                sSub[v * SS + task] = S0(v, ii, jj) - dt / ldx * (fx[ii + (s + 1) * (jj - 1)] - fx[(ii - 1) + (s + 1) * (jj - 1)]) -
                                      dt / ldy * (fy[(ii - 1) + s * jj] - fy[(ii - 1) + s * (jj - 1)]);
            }
        }
#undef S0
        __syncthreads();
        subcell_finish<P, N, 64>(m, b, gcell, slot, sa.lh0, sa.lhcap, par, flag, sSub, sUh, sLim, sLob, sDt, sRng,
                                 &sFlag, &sNth, o);
    }
}

// ---------------------------------------------------------------------------
// Baer-Nunziato subcell_recompute over the compacted list, one wavefront per
// cell (Baer-Nunziato/limiter.f90:416-596): slopes also on the halo ring, so
// the subgrid spans -1..s+2 and takes all 8 neighbours; NCP source in the
// half step and the update; Roe fluctuations at the subcell faces; no
// boundary overrides.  The half-step states w are recomputed from (S0,
// slopes, Qt) where needed instead of being stored.
// ---------------------------------------------------------------------------
// Baer-Nunziato subcell_recompute over the compacted list (Baer-Nunziato/limiter.f90:416-596),
// one workgroup of kSubBnBlock threads per listed cell.  The reference evaluates the
// path-conservative RusanovFluxS of every interior subcell face twice (as the right face of one
// subcell and the left face of the next, with the same arguments); here each face is evaluated
// once, one face per thread, and the subcell updates read both sides from LDS: the same values,
// with half the flux work and no serial loop over a subcell's four faces.
template <int N>
__global__ __launch_bounds__(kSubBnBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_subcell_bn(StepBufs b)
{
    using P = BnPDE;
    using S = Shape<P, N>;
    constexpr int NT = kSubBnBlock;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, V = S::V;
    constexpr int G = s + 4, GG = G * G;  // subgrid index -1..s+2
    constexpr int H = s + 2, HH = H * H;  // halo ring 0..s+1
    constexpr int NF = (s + 1) * s;       // x-faces (and y-faces) of the interior subcells
    constexpr int FP = (2 * NF + NT - 1) / NT;  // faces per thread
    constexpr int LY = 9 * V * n * s, SL = 3 * V * HH, FL = 4 * V * NF;
    constexpr int DMPB = V * SS + V * NN;
    constexpr int R0 = cmax(cmax(LY, SL), cmax(FL, DMPB));
    __shared__ double sS0[V * GG];
    __shared__ double sR0[R0];     // limy of 9 blocks; slopes x / y / t; face F* and D; DMP scratch
    __shared__ double sSub[V * SS];  // NCP source of the interior half-step states, then the new averages
    __shared__ double sUh[V * NN];
    __shared__ double sDt[NN];
    __shared__ double sRng[8 * V + 1];
    __shared__ int sFlag, sNth;
    __shared__ int sNb[9], sLs[9];  // the 3x3 neighbourhood: group cell and pool slot (-1 unlimited)
    const OpPtr op = (OpPtr)b.op;
    const int lane = threadIdx.x;
    const int count = b.counters[kCntList];
    double *sLy = sR0;
    double *slx = sR0, *sly = sR0 + V * HH, *slt = sR0 + 2 * V * HH;
    double *sFF = sR0, *sFD = sR0 + 2 * V * NF;  // [dir][v][face]: F* and D of x- (dir 0) and y-faces
    __shared__ SubOpLds<N> sOps;
    const SubOps o = stage_subops<N>(op, sOps, lane, NT);
    for (int item = blockIdx.x; item < count; item += gridDim.x) {
        const int gcell = b.list[item];
        const int slot = ld_const(b.cell_solve + gcell);
        const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + slot);
        const auto &m = sa.m;
        const SolveState st = ld_const(b.state + slot);
        const int par = st.parity;
        const double dt = st.dt;
        const double ldx = m.dx / (double)s, ldy = m.dy / (double)s;
        const int cell = gcell - sa.cell0;  // local cell of the solve
        const int flag = b.rc_final[gcell];
        const int ci = cell % m.imax, cj = cell / m.imax;
        // the neighbourhood's cells and limiter slots once per cell, status and slot loaded
        // side by side (one memory round trip, not two per task of the passes below)
        if (lane < 9) {
            const int di = lane % 3 - 1, dj = lane / 3 - 1;
            const int nb = sa.cell0 + nb_index(ci, di, m.imax, m.wrap & 1) + m.imax * nb_index(cj, dj, m.jmax, m.wrap >> 1);
            const int stt = b.status[par][nb], ls = b.lh_slot[par][nb];
            sNb[lane] = nb;
            sLs[lane] = stt != 0 ? ls : -1;  // lim_slot
        }
        __syncthreads();
        // blocks (di,dj) of the 3x3 Voronoi neighbourhood, :463-479: a thread loads the n x V
        // values of column a of a neighbour once and forms half of that column's limy rows qq
        // (one round of loads for the workgroup instead of one per qq)
        constexpr int QH = (s + 1) / 2;
        for (int task = lane; task < 9 * n * 2; task += NT) {
            const int blk = task / (2 * n), r = task - blk * 2 * n, a = r >> 1, q0 = (r & 1) * QH;
            if (sLs[blk] >= 0) continue;
            const double *u = b.uh[par] + (size_t)sNb[blk] * V * NN;
            double x[n][V];
#pragma unroll
            for (int bb = 0; bb < n; bb++)
#pragma unroll
                for (int v = 0; v < V; v++) x[bb][v] = u[v + V * (a + n * bb)];
#pragma unroll
            for (int dq = 0; dq < QH; dq++) {
                const int qq = q0 + dq;
                if (qq >= s) break;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double acc = 0.0;
#pragma unroll
                    for (int bb = 0; bb < n; bb++) acc = acc + x[bb][v] * o.u2l[qq + s * bb];
                    sLy[(blk * V + v) * n * s + a + n * qq] = acc;
                }
            }
        }
        __syncthreads();
        for (int task = lane; task < GG; task += NT) {  // subuh0 on -1..s+2
            const int ii = task % G - 1, jj = task / G - 1;
            const int di = ii < 1 ? -1 : (ii > s ? 1 : 0), dj = jj < 1 ? -1 : (jj > s ? 1 : 0);
            const int p = ii - 1 - di * s, qq = jj - 1 - dj * s;
            const int blk = (di + 1) + 3 * (dj + 1);
            const int ls = sLs[blk];
            if (ls < 0) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < n; a++) acc = acc + sLy[(blk * V + v) * n * s + a + n * qq] * o.u2l[p + s * a];
                    sS0[v * GG + task] = acc;
                }
            } else {
                const double *L = b.lh[par] + (size_t)ls * V * SS;
#pragma unroll
                for (int v = 0; v < V; v++) sS0[v * GG + task] = L[v + V * (p + s * qq)];
            }
            // the side codes (inflow / reflective) on ghosts outside the domain
            const int gx = (ci == 0 && di < 0) ? -1 : ((ci == m.imax - 1 && di > 0) ? 1 : 0);
            const int gy = (cj == 0 && dj < 0) ? -1 : ((cj == m.jmax - 1 && dj > 0) ? 1 : 0);
            if (gx || gy) side_subcell<P>(m, gx, gy, sS0 + task, GG);
        }
        __syncthreads();
#define S0(v, ii, jj) sS0[(v) * GG + ((ii) + 1) + G * ((jj) + 1)]
#define HIX(ii, jj) ((ii) + H * (jj))
        // slopes, NCP source and half-step time derivative on 0..s+1, :486-513; for the
        // interior subcells also the NCP source of the half-step state (:520-523)
        for (int task = lane; task < HH; task += NT) {
            const int ii = task % H, jj = task / H;
            double sx[V], sy[V], Gr[2 * V], src[V], q0[V], wL[V], wR[V], FL[2 * V], FR[2 * V], Qt[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                sx[v] = minmod(S0(v, ii + 1, jj) - S0(v, ii, jj), S0(v, ii, jj) - S0(v, ii - 1, jj));
                sy[v] = minmod(S0(v, ii, jj + 1) - S0(v, ii, jj), S0(v, ii, jj) - S0(v, ii, jj - 1));
                Gr[v] = sx[v] / ldx;
                Gr[V + v] = sy[v] / ldy;
                q0[v] = S0(v, ii, jj);
            }
            P::ncp(m, q0, Gr, src);
#pragma unroll
            for (int v = 0; v < V; v++) {
                wL[v] = q0[v] - 0.5 * sx[v];
                wR[v] = q0[v] + 0.5 * sx[v];
            }
            P::flux(m, wL, FL);
            P::flux(m, wR, FR);
#pragma unroll
            for (int v = 0; v < V; v++) Qt[v] = src[v] - (FR[v] - FL[v]) / ldx;
#pragma unroll
            for (int v = 0; v < V; v++) {
                wL[v] = q0[v] - 0.5 * sy[v];
                wR[v] = q0[v] + 0.5 * sy[v];
            }
            P::flux(m, wL, FL);
            P::flux(m, wR, FR);
#pragma unroll
            for (int v = 0; v < V; v++) {
                Qt[v] = Qt[v] - (FR[V + v] - FL[V + v]) / ldy;
                slx[v * HH + task] = sx[v];
                sly[v * HH + task] = sy[v];
                slt[v * HH + task] = Qt[v];
            }
            if (ii >= 1 && ii <= s && jj >= 1 && jj <= s) {
#pragma unroll
                for (int v = 0; v < V; v++) q0[v] = q0[v] + 0.5 * dt * Qt[v];
                P::ncp(m, q0, Gr, src);
#pragma unroll
                for (int v = 0; v < V; v++) sSub[v * SS + (ii - 1) + s * (jj - 1)] = src[v];
            }
        }
        __syncthreads();
        // half-step face states: w = (S0 -/+ 0.5 slope) + 0.5 dt Qt, as the reference stores them
        auto wstate = [&](int v, int ii, int jj, int dir, int right) {
            const double sl = dir == 0 ? slx[v * HH + HIX(ii, jj)] : sly[v * HH + HIX(ii, jj)];
            const double w0 = right ? S0(v, ii, jj) + 0.5 * sl : S0(v, ii, jj) - 0.5 * sl;
            return w0 + 0.5 * dt * slt[v * HH + HIX(ii, jj)];
        };
        // RusanovFluxS of every interior face, Baer-Nunziato/limiter.f90:359-383: x-face r joins
        // subcells (r % (s+1), row) and (r % (s+1) + 1, row), row = r / (s+1) + 1; y-face r joins
        // (col, r / s) and (col, r / s + 1), col = r % s + 1
        double fF[FP][V], fD[FP][V];
#pragma unroll
        for (int k = 0; k < FP; k++) {
            const int task = lane + k * NT;
            if (task >= 2 * NF) continue;
            const int dir = task / NF, r = task - dir * NF;
            int iL, jL, iR, jR;
            if (dir == 0) {
                iL = r % (s + 1);
                jL = r / (s + 1) + 1;
                iR = iL + 1;
                jR = jL;
            } else {
                iL = r % s + 1;
                jL = r / s;
                iR = iL;
                jR = jL + 1;
            }
            const double nx = dir == 0 ? 1.0 : 0.0, ny = dir == 0 ? 0.0 : 1.0;
            double QL[V], QR[V], FLf[2 * V], FRf[2 * V], B9[V], dQ[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                QL[v] = wstate(v, iL, jL, dir, 1);
                QR[v] = wstate(v, iR, jR, dir, 0);
            }
            const double smax = f_max(P::maxabs_eig(m, QL, nx, ny), P::maxabs_eig(m, QR, nx, ny));
            P::flux(m, QL, FLf);
            P::flux(m, QR, FRf);
            P::roe_col9(m, QL, QR, nx, ny, B9);
#pragma unroll
            for (int v = 0; v < V; v++) dQ[v] = QR[v] - QL[v];
#pragma unroll
            for (int v = 0; v < V; v++) {
                double mm = 0.0;
                mm = mm + (FLf[v] + FRf[v]) * nx;
                mm = mm + (FLf[V + v] + FRf[V + v]) * ny;
                fF[k][v] = 0.5 * mm - 0.5 * smax * dQ[v];
                fD[k][v] = P::roe_times(B9, v, dQ);
            }
        }
        __syncthreads();  // the slopes are read: their space takes the face values
#pragma unroll
        for (int k = 0; k < FP; k++) {
            const int task = lane + k * NT;
            if (task >= 2 * NF) continue;
            const int dir = task / NF, r = task - dir * NF;
#pragma unroll
            for (int v = 0; v < V; v++) {
                sFF[(dir * V + v) * NF + r] = fF[k][v];
                sFD[(dir * V + v) * NF + r] = fD[k][v];
            }
        }
        __syncthreads();
        // update of the interior subcells, :516-537
        for (int task = lane; task < SS; task += NT) {
            const int ii = task % s + 1, jj = task / s + 1;
            const int xr = ii + (s + 1) * (jj - 1), xl = xr - 1;      // right / left x-face
            const int yt = (ii - 1) + s * jj, yb = yt - s;            // top / bottom y-face
#pragma unroll
            for (int v = 0; v < V; v++) {
                const double Fp_ = sFF[v * NF + xr], Fm_ = sFF[v * NF + xl];
                const double Gp_ = sFF[(V + v) * NF + yt], Gm_ = sFF[(V + v) * NF + yb];
                const double Dpx = sFD[v * NF + xr], Dmx = sFD[v * NF + xl];
                const double Dpy = sFD[(V + v) * NF + yt], Dmy = sFD[(V + v) * NF + yb];
                const double src = sSub[v * SS + task];
                sSub[v * SS + task] = S0(v, ii, jj) - dt / ldx * (Fp_ - Fm_) - dt / ldy * (Gp_ - Gm_) -
                                      dt / ldx * 0.5 * (Dpx + Dmx) - dt / ldy * 0.5 * (Dpy + Dmy) + dt * src;
            }
        }
#undef S0
#undef HIX
        __syncthreads();
        double *sLim = sR0, *sLob = sLim + V * SS;
        subcell_finish<P, N, NT>(m, b, gcell, slot, sa.lh0, sa.lhcap, par, flag, sSub, sUh, sLim, sLob, sDt, sRng,
                                 &sFlag, &sNth, o);
    }
}

// calc_time_step (Euler/calc_time_step.f90:9-52) for every active solve: each
// workgroup reduces the candidates of a piece of one solve and folds the result
// into the solve's minimum with an integer atomic on the bit pattern (positive
// doubles order like their bits; NaN candidates are skipped as MINVAL skips
// them); the last workgroup to finish turns the minima into dt.  No host sync,
// no memset.
// get_min_max step 2, Euler/limiter.f90:147-167: the DMP bounds of every cell, the min / max over
// its 3x3 neighbourhood of the first-pass ranges (lmin_old / lmax_old, left by the previous
// step's update and subcell passes or by the seed).  One thread per cell, consecutive cells on
// consecutive lanes: each lane loads its own column of the three rows (coalesced), and takes
// the left / right columns from the adjacent lanes (ds_bpermute) wherever they hold the
// adjacent cells of the same row; the lanes at a wave's ends and at row ends (periodic or
// clamped neighbours) load theirs.  The values are folded in the reference order (own value,
// then the nine neighbours row by row).  Six loads per variable instead of twenty (Baer-Nunziato
// ensemble, N=3 arena: 0.63 -> 0.59 ms).  In the predictor kernels this gather sat at the end of
// each entry with its latency exposed.
template <class P>
__global__ __launch_bounds__(kBlock) void k_ranges(StepBufs b)
{
    const int ntab = ld_const(b.meta + kTabCell);
    const int lane = threadIdx.x & 63;
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = ld_const(b.tab[kTabCell] + e);
        if (!ld_const(b.state + w.x).active) continue;
        const ADG_CONST SolveArgs *sa = (const ADG_CONST SolveArgs *)(b.solves + w.x);
        const MeshArgs m = sa->m;
        // every lane takes part in the exchanges: lanes past the last cell shadow it
        const bool live = w.y + (int)threadIdx.x < m.ncell;
        const int cell = live ? w.y + (int)threadIdx.x : m.ncell - 1;
        const int ci = cell % m.imax, cj = cell / m.imax;
        const int il = nb_index(ci, -1, m.imax, m.wrap & 1), ir = nb_index(ci, 1, m.imax, m.wrap & 1);
        const bool from_l = lane > 0 && il == ci - 1;   // lane - 1 holds cell - 1 of this row
        const bool from_r = lane < 63 && ir == ci + 1;  // lane + 1 holds cell + 1 (then below ncell)
        const int al = (lane - 1) << 2, ar = (lane + 1) << 2;
        int rows[3];
#pragma unroll
        for (int jj = -1; jj <= 1; jj++) rows[jj + 1] = m.imax * nb_index(cj, jj, m.jmax, m.wrap >> 1);
#pragma unroll
        for (int v = 0; v < P::V; v++) {
            const size_t row = (size_t)v * b.G + sa->cell0;
            double cm[3], cx[3], lm[3], lx[3], rm[3], rx[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                cm[r] = b.lmin_old[row + ci + rows[r]];
                cx[r] = b.lmax_old[row + ci + rows[r]];
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
                lm[r] = lane_get(cm[r], al);
                lx[r] = lane_get(cx[r], al);
                rm[r] = lane_get(cm[r], ar);
                rx[r] = lane_get(cx[r], ar);
            }
            if (!from_l) {
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    lm[r] = b.lmin_old[row + il + rows[r]];
                    lx[r] = b.lmax_old[row + il + rows[r]];
                }
            }
            if (!from_r) {
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    rm[r] = b.lmin_old[row + ir + rows[r]];
                    rx[r] = b.lmax_old[row + ir + rows[r]];
                }
            }
            double mn = cm[1], mx = cx[1];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                mn = f_min(mn, lm[r]);
                mx = f_max(mx, lx[r]);
                mn = f_min(mn, cm[r]);
                mx = f_max(mx, cx[r]);
                mn = f_min(mn, rm[r]);
                mx = f_max(mx, rx[r]);
            }
            if (live) {
                b.lmin[row + cell] = mn;
                b.lmax[row + cell] = mx;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_dt(StepBufs b, double host_time, double host_tend, int use_host_time)
{
    __shared__ double red[kBlock];
    __shared__ int last;
    const int tid = threadIdx.x;
    const int ntab = b.meta[kTabCell];
    for (int e = xcd_entry(blockIdx.x, gridDim.x); e < ntab; e += gridDim.x) {
        const int2 w = b.tab[kTabCell][e];
        if (!b.state[w.x].active) continue;
        const SolveArgs &sa = b.solves[w.x];
        const int cell = w.y + tid;
        red[tid] = cell < sa.m.ncell ? b.cell_dt[sa.cell0 + cell] : __builtin_nan("");
        __syncthreads();
        for (int k = kBlock / 2; k > 0; k >>= 1) {
            if (tid < k) red[tid] = fold_min(red[tid], red[tid + k]);
            __syncthreads();
        }
        if (tid == 0 && !__builtin_isnan(red[0]))
            atomicMin(&b.state[w.x].dtmin, (unsigned long long)__double_as_longlong(red[0]));
        __syncthreads();
    }
    if (tid == 0) {
        __threadfence();
        const int ticket = atomicAdd(&b.counters[kCntDtBlocks], 1);
        last = ticket == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int ns = b.meta[kMetaSolves];
    for (int k = tid; k < ns; k += kBlock) {
        SolveState &st = b.state[k];
        if (!st.active) continue;
        const unsigned long long bits = __hip_atomic_load(&st.dtmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double mn = bits == ~0ull ? __builtin_nan("") : __longlong_as_double((long long)bits);
        const double time = use_host_time ? host_time : st.time;
        const double tend = use_host_time ? host_tend : st.tend;
        double dt = f_min(1.0e20, mn);           // dt = 1e20 then min over nodes (:257-269)
        if (time + dt > tend) dt = tend - time;  // :276-278
        st.dt = dt;
        st.time = time;
        st.tend = tend;
        st.dtmin = ~0ull;
    }
    if (tid == 0) b.counters[kCntDtBlocks] = 0;
}

// Per-solve bookkeeping after a step: kPartAdvance (time = time + dt, count the
// step, switch buffers, end the solve at its step count or end time),
// kPartFlip (switch buffers only: update_limiter, Euler/limiter.f90:272-296),
// kPartSeedDone (the seeds of freshly uploaded solutions are in place).  All three
// also end the step's count of limited-cell pool slots.
__global__ __launch_bounds__(kBlock) void k_solves(StepBufs b, int part)
{
    const int ns = b.meta[kMetaSolves];
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < ns; k += gridDim.x * kBlock) {
        SolveState &st = b.state[k];
        // limited-cell pool: the slots the step (or the seeding) took; more than the solve's
        // segment holds ends the solve as failed (its last solution is whole, the subcell
        // averages of some limited cells are not kept); the count restarts for the next step
        const int taken = b.lh_count[k];
        b.lh_count[k] = 0;
        const bool over = taken > ld_const(b.solves + k).lhcap;
        if (part == kPartSeedDone) {
            st.seed = 0;
            if (over) {
                st.failed = 1;
                st.active = 0;
            }
            continue;
        }
        if (!st.active) continue;
        st.parity ^= 1;
        if (part == kPartAdvance) {
            st.time = st.time + st.dt;
            st.steps += 1;
            if (st.steps >= st.max_steps || (st.stop_at_tend && !(st.time < st.tend))) st.active = 0;
        }
        if (over) {
            st.failed = 1;
            st.active = 0;
        }
    }
}

// Work tables from per-table prefix counts over the solve slots: entry e of
// table k belongs to the last slot s with prefix[s] <= e.
__global__ __launch_bounds__(kBlock) void k_tables(StepBufs b, const int *prefix, int nslot, int4 quanta)
{
    const int q[kNumTabs] = {quanta.x, quanta.y, quanta.z, quanta.w};
    int *meta = const_cast<int *>(b.meta);
    for (int k = 0; k < kNumTabs; k++) {
        const int *pk = prefix + k * (nslot + 1);
        const int total = pk[nslot];
        int2 *tab = const_cast<int2 *>(b.tab[k]);
        for (int e = blockIdx.x * kBlock + threadIdx.x; e < total; e += gridDim.x * kBlock) {
            int lo = 0, hi = nslot - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pk[mid] <= e)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            tab[e] = make_int2(lo, (e - pk[lo]) * q[k]);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) meta[k] = total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) meta[kMetaSolves] = nslot;
}

// Initial state of one solve from a formula: one thread per (cell, node).
// ---------------------------------------------------------------------------
// Troubled-cell seeding of a formula initial condition, the end of initialize
// (Euler/initialize.f90:274-303, Baer-Nunziato/initialize.f90:272-301): a cell
// whose projected state fails the DMP against its own nodal range with the
// relaxed margin 0.1 (DMP(..., 1.0e-1), Euler/limiter.f90 DMP), or the positivity
// check of its subcell values, starts limited, with its subcell averages taken
// from the formula at the subcell barycentres corner + xilimbary * dx
// (xilimbary(i) = 0.5/nSubLim + (i-1)/nSubLim, Euler/initialize.f90:211-213).  The
// reference writes x0 for the corner without setting it; the corner is meant.
// One 64-thread workgroup per cell of solve `slot`, solution in uh[parity].
// ---------------------------------------------------------------------------
template <class P, int N>
__global__ __launch_bounds__(64) void k_seed_ic(StepBufs b, IcSpec ic, int slot, int ncell)
{
    using S = Shape<P, N>;
    constexpr int n = S::n, NN = S::NN, s = S::s, SS = S::SS, V = S::V;
    __shared__ double sUh[V * NN], sTy[V * n * s], sBy[V * NN], sLim[V * SS], sLob[V * NN];
    __shared__ int sFail;
    const OpPtr op = (OpPtr)b.op;
    const int lane = threadIdx.x;
    const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + slot);
    const auto &m = sa.m;
    const int par = ld_const(b.state + slot).parity;
    for (int cell = blockIdx.x; cell < ncell; cell += gridDim.x) {
        const size_t g = (size_t)sa.cell0 + cell;
        const double *u = b.uh[par] + g * V * NN;
        for (int k = lane; k < V * NN; k += 64) {
            const int v = k / NN, node = k - v * NN;
            sUh[v * NN + node] = u[v + V * node];
        }
        if (lane == 0) sFail = 0;
        __syncthreads();
        cell_ranges_stage1<P, N>(op->u2l, op->u2b, sUh, sTy, sBy, lane, 64);
        __syncthreads();
        if (cell_ranges_stage2<P, N>(op->u2l, op->u2b, m, sTy, sBy, sLim, sLob, lane, 64)) atomicOr(&sFail, 1);
        __syncthreads();
        for (int v = lane; v < V; v += 64) {
            double umn = __builtin_nan(""), umx = umn, smn = umn, smx = umn, bmn = umn, bmx = umn;
            for (int k = 0; k < NN; k++) {
                umn = fold_min(umn, sUh[v * NN + k]);
                umx = fold_max(umx, sUh[v * NN + k]);
                bmn = fold_min(bmn, sLob[v * NN + k]);
                bmx = fold_max(bmx, sLob[v * NN + k]);
            }
            for (int k = 0; k < SS; k++) {
                smn = fold_min(smn, sLim[v * SS + k]);
                smx = fold_max(smx, sLim[v * SS + k]);
            }
            // lmin = min(minval(uh), minval(lim), minval(lob)) against the nodal range, margin max(1e-4, 1e-3 range, 0.1)
            const double mn = f_min(f_min(umn, smn), bmn), mx = f_max(f_max(umx, smx), bmx);
            const double ldiff = f_max(f_max(1.0e-4, 1.0e-3 * (umx - umn)), 1.0e-1);
            if (mn < umn - ldiff || mx > umx + ldiff) atomicOr(&sFail, 2);
        }
        __syncthreads();
        if (sFail) {
            int nth = 0;
            if (lane == 0) nth = atomicAdd(b.lh_count + slot, 1);
            nth = __shfl(nth, 0);
            const int ls = nth < sa.lhcap ? sa.lh0 + nth : -1;  // none left: kPartSeedDone fails the solve
            if (lane == 0) {
                b.status[par][g] = 1;
                b.lh_slot[par][g] = ls;
            }
            const int ci = cell % m.imax, cj = cell / m.imax;
            const double cx = (m.xl[0] + ((double)ci + 0.5) * m.dx) - 0.5 * m.dx;
            const double cy = (m.xl[1] + ((double)cj + 0.5) * m.dy) - 0.5 * m.dy;
            for (int k = lane; k < SS; k += 64) {
                const int ii = k % s, jj = k / s;
                const double bx = 0.0 + 0.5 / (double)s + (double)ii * 1.0 / (double)s;
                const double by = 0.0 + 0.5 / (double)s + (double)jj * 1.0 / (double)s;
                double Q[V];
                ic_field(ic, m.gamma, m.eos, cx + bx * m.dx, cy + by * m.dy, Q);
#pragma unroll
                for (int v = 0; v < V; v++)
                    if (ls >= 0) b.lh[par][(size_t)ls * V * SS + v + V * k] = Q[v];
            }
        }
        __syncthreads();
    }
}

template <class P, int N>
__global__ __launch_bounds__(kBlock) void k_init(StepBufs b, IcSpec ic, int slot, int parity)
{
#pragma clang fp contract(off)
    constexpr int V = P::V, n = N + 1, NN = n * n;
    const ADG_CONST SolveArgs &sa = *(const ADG_CONST SolveArgs *)(b.solves + slot);
    const auto &m = sa.m;
    const OpPtr op = (OpPtr)b.op;
    const long t = blockIdx.x * (long)kBlock + threadIdx.x;
    if (t >= (long)m.ncell * NN) return;
    const int cell = (int)(t / NN), node = (int)(t - (long)cell * NN);
    const int i = cell % m.imax, j = cell / m.imax, k = node % n, l = node / n;
    const double cx = (m.xl[0] + ((double)i + 0.5) * m.dx) - 0.5 * m.dx;  // Euler/initialize.f90:151,265
    const double cy = (m.xl[1] + ((double)j + 0.5) * m.dy) - 0.5 * m.dy;
    double Q[V];
    ic_field(ic, m.gamma, m.eos, cx + op->xi[k] * m.dx, cy + op->xi[l] * m.dy, Q);
    double *u = b.uh[parity] + ((size_t)sa.cell0 + cell) * V * NN + V * node;
#pragma unroll
    for (int v = 0; v < V; v++) u[v] = Q[v];
}

template <int N>
constexpr bool node_variant_fits(int variant)
{
    return variant == kPredNode64 ? (N + 1) * (N + 1) * (N + 1) <= 64
                                  : (variant == kPredNode128 ? (N + 1) * (N + 1) * (N + 1) <= 128 : true);
}

template <class P, int N>
hipError_t launch_predictor(const StepBufs &b, hipStream_t st, int variant, int grid)
{
    using S = Shape<P, N>;
    switch (variant) {
    case kPredLine:
        hipLaunchKernelGGL((k_predictor<P, N>), dim3(grid), dim3(S::PRED_BLOCK), 0, st, b);
        break;
    case kPredNode64:
        if constexpr ((N + 1) * (N + 1) * (N + 1) <= 64)
            hipLaunchKernelGGL((k_predictor_nodes<P, N, 64>), dim3(grid), dim3(64), 0, st, b);
        else
            return hipErrorInvalidValue;
        break;
    case kPredNode128:
        hipLaunchKernelGGL((k_predictor_nodes<P, N, 128>), dim3(grid), dim3(128), 0, st, b);
        break;
    case kPredNode256:
        hipLaunchKernelGGL((k_predictor_nodes<P, N, 256>), dim3(grid), dim3(256), 0, st, b);
        break;
    case kPredLean:
        hipLaunchKernelGGL((k_predictor_lean<P, N>), dim3(grid), dim3(S::PRED_BLOCK), 0, st, b);
        break;
    case kPredMfma:
        if constexpr (kMfmaBuild && N <= 3)
            hipLaunchKernelGGL((k_predictor_mfma<P, N>), dim3(grid), dim3(kMfBlock<P>), 0, st, b);
        else
            return hipErrorInvalidValue;
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipSuccess;
}

template <class P, int N>
hipError_t launch_step_impl(const StepBufs &b, hipStream_t st, int part, const int grid[kNumTabs], int variant,
                            int upd)
{
    using S = Shape<P, N>;
    switch (part) {
    case kPartRanges:
        hipLaunchKernelGGL((k_ranges<P>), dim3(grid[kTabCell]), dim3(kBlock), 0, st, b);
        break;
    case kPartPredictor: {
        const hipError_t e = launch_predictor<P, N>(b, st, variant, grid[kTabPred]);
        if (e != hipSuccess) return e;
        break;
    }
    case kPartFace:
        hipLaunchKernelGGL((k_face<P, N>), dim3(grid[kTabFace]), dim3(kBlock), 0, st, b);
        break;
    case kPartUpdate:
        if (upd == kUpdShared)
            hipLaunchKernelGGL((k_update_shared<P, N, false>), dim3(grid[kTabUpd]), dim3(S::UPD_BLOCK), 0, st, b);
        else
            hipLaunchKernelGGL((k_update<P, N, false>), dim3(grid[kTabUpd]), dim3(S::UPD_BLOCK), 0, st, b);
        break;
    case kPartSeed:
        if (upd == kUpdShared)
            hipLaunchKernelGGL((k_update_shared<P, N, true>), dim3(grid[kTabUpd]), dim3(S::UPD_BLOCK), 0, st, b);
        else
            hipLaunchKernelGGL((k_update<P, N, true>), dim3(grid[kTabUpd]), dim3(S::UPD_BLOCK), 0, st, b);
        break;
    case kPartSpread:
        hipLaunchKernelGGL(k_spread, dim3(grid[kTabCell]), dim3(kBlock), 0, st, b);
        break;
    case kPartSubcell: {
        const int g = b.G < kSubcellGrid ? b.G : kSubcellGrid;
        if (P::kNcp)
            hipLaunchKernelGGL((k_subcell_bn<N>), dim3(g), dim3(kSubBnBlock), 0, st, b);
        else
            hipLaunchKernelGGL((k_subcell<N>), dim3(g), dim3(64), 0, st, b);
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_part_pde(int N, const StepBufs &b, hipStream_t st, int part, const int grid[kNumTabs], int variant,
                           int upd)
{
    switch (N) {
    case 1: return launch_step_impl<P, 1>(b, st, part, grid, variant, upd);
    case 2: return launch_step_impl<P, 2>(b, st, part, grid, variant, upd);
    case 3: return launch_step_impl<P, 3>(b, st, part, grid, variant, upd);
    case 4: return launch_step_impl<P, 4>(b, st, part, grid, variant, upd);
    default: return hipErrorInvalidValue;
    }
}

template <class P, int N>
hipError_t launch_test_impl(const MeshArgs &m, const StepBufs &b, const TestBufs &t, double dt, hipStream_t st,
                            int variant)
{
    using S = Shape<P, N>;
    constexpr int N3 = (N + 1) * (N + 1) * (N + 1);
    switch (variant) {
    case kPredLine: {
        const int g = (m.ncell + S::CPW - 1) / S::CPW;
        hipLaunchKernelGGL((k_predictor_test<P, N>), dim3(g), dim3(S::PRED_BLOCK), 0, st, m, b, t, dt);
        break;
    }
    case kPredNode64:
        if constexpr (N3 <= 64) {
            const int g = (m.ncell + 64 / N3 - 1) / (64 / N3);
            hipLaunchKernelGGL((k_predictor_nodes_test<P, N, 64>), dim3(g), dim3(64), 0, st, m, b, t, dt);
        } else {
            return hipErrorInvalidValue;
        }
        break;
    case kPredNode128: {
        const int g = (m.ncell + 128 / N3 - 1) / (128 / N3);
        hipLaunchKernelGGL((k_predictor_nodes_test<P, N, 128>), dim3(g), dim3(128), 0, st, m, b, t, dt);
        break;
    }
    case kPredNode256: {
        const int g = (m.ncell + 256 / N3 - 1) / (256 / N3);
        hipLaunchKernelGGL((k_predictor_nodes_test<P, N, 256>), dim3(g), dim3(256), 0, st, m, b, t, dt);
        break;
    }
    case kPredLean: {
        const int g = (m.ncell + S::CPW - 1) / S::CPW;
        hipLaunchKernelGGL((k_predictor_lean_test<P, N>), dim3(g), dim3(S::PRED_BLOCK), 0, st, m, b, t, dt);
        break;
    }
    case kPredMfma:
        if constexpr (kMfmaBuild && N <= 3) {
            const int g = (m.ncell + MfShape<P, N>::CPW - 1) / MfShape<P, N>::CPW;
            hipLaunchKernelGGL((k_predictor_mfma_test<P, N>), dim3(g), dim3(kMfBlock<P>), 0, st, m, b, t, dt);
        } else {
            return hipErrorInvalidValue;
        }
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_test_pde(int N, const MeshArgs &m, const StepBufs &b, const TestBufs &t, double dt, hipStream_t st,
                           int variant)
{
    switch (N) {
    case 1: return launch_test_impl<P, 1>(m, b, t, dt, st, variant);
    case 2: return launch_test_impl<P, 2>(m, b, t, dt, st, variant);
    case 3: return launch_test_impl<P, 3>(m, b, t, dt, st, variant);
    case 4: return launch_test_impl<P, 4>(m, b, t, dt, st, variant);
    default: return hipErrorInvalidValue;
    }
}

template <class P, int N>
void quanta_impl(int variant, int q[kNumTabs])
{
    constexpr int N3 = (N + 1) * (N + 1) * (N + 1);
    if constexpr (N <= 3) {
        if (variant == kPredMfma) {
            q[kTabPred] = MfShape<P, N>::CPW;
            q[kTabUpd] = Shape<P, N>::CPWU;
            return;
        }
    }
    q[kTabPred] = (variant == kPredLine || variant == kPredLean) ? Shape<P, N>::CPW
                  : variant == kPredNode64 ? 64 / N3 : variant == kPredNode128 ? 128 / N3 : 256 / N3;
    q[kTabUpd] = Shape<P, N>::CPWU;
}

template <class P>
void quanta_pde(int N, int variant, int q[kNumTabs])
{
    switch (N) {
    case 1: quanta_impl<P, 1>(variant, q); break;
    case 2: quanta_impl<P, 2>(variant, q); break;
    case 3: quanta_impl<P, 3>(variant, q); break;
    default: quanta_impl<P, 4>(variant, q); break;
    }
    q[kTabFace] = kBlock;
    q[kTabCell] = kBlock;
}

}  // namespace

bool pred_variant_valid(int pde, int N, int variant)
{
    if (N < 1 || N > 4) return false;
    if (pde == kPdeAd) return variant == kPredLine;
    if (variant == kPredMfma) return kMfmaBuild && N <= 3;
    if (variant == kPredLine || variant == kPredNode128 || variant == kPredNode256 || variant == kPredLean) return true;
    return variant == kPredNode64 && (N + 1) * (N + 1) * (N + 1) <= 64;
}

int pred_cells_per_block(int pde, int N, int variant)
{
    int q[kNumTabs];
    table_quanta(pde, N, variant, q);
    return q[kTabPred];
}

namespace {

template <class F>
int resident_of(F kernel, int block)
{
    int nb = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) == hipSuccess ? nb : 0;
}

template <class P, int N>
int resident_impl(int variant, int upd, int tab)
{
    using S = Shape<P, N>;
    switch (tab) {
    case kTabPred:
        switch (variant) {
        case kPredLine: return resident_of(k_predictor<P, N>, S::PRED_BLOCK);
        case kPredNode64:
            if constexpr ((N + 1) * (N + 1) * (N + 1) <= 64) return resident_of(k_predictor_nodes<P, N, 64>, 64);
            return 0;
        case kPredNode128: return resident_of(k_predictor_nodes<P, N, 128>, 128);
        case kPredNode256: return resident_of(k_predictor_nodes<P, N, 256>, 256);
        case kPredLean: return resident_of(k_predictor_lean<P, N>, S::PRED_BLOCK);
        case kPredMfma:
            if constexpr (kMfmaBuild && N <= 3) return resident_of(k_predictor_mfma<P, N>, kMfBlock<P>);
            return 0;
        default: return 0;
        }
    case kTabUpd:
        return upd == kUpdShared ? resident_of(k_update_shared<P, N, false>, S::UPD_BLOCK)
                                 : resident_of(k_update<P, N, false>, S::UPD_BLOCK);
    case kTabFace: return resident_of(k_face<P, N>, kBlock);
    default: return 0;
    }
}

template <class P>
int resident_pde(int N, int variant, int upd, int tab)
{
    switch (N) {
    case 1: return resident_impl<P, 1>(variant, upd, tab);
    case 2: return resident_impl<P, 2>(variant, upd, tab);
    case 3: return resident_impl<P, 3>(variant, upd, tab);
    case 4: return resident_impl<P, 4>(variant, upd, tab);
    default: return 0;
    }
}

}  // namespace

int resident_blocks_per_cu(int pde, int N, int variant, int upd, int tab)
{
    if (pde == kPdeBn) return resident_pde<BnPDE>(N, variant, upd, tab);
    if (pde == kPdeEuler) return resident_pde<EulerPDE>(N, variant, upd, tab);
    return 0;
}

void table_quanta(int pde, int N, int variant, int quanta[kNumTabs])
{
    if (pde == kPdeAd)
        table_quanta_ad(N, quanta);
    else if (pde == kPdeBn)
        quanta_pde<BnPDE>(N, variant, quanta);
    else
        quanta_pde<EulerPDE>(N, variant, quanta);
}

hipError_t launch_part(int pde, int N, int variant, int upd, const StepBufs &b, hipStream_t st, int part,
                       const int grid[kNumTabs], int nslot_cap)
{
    if (part == kPartDt) return launch_dt(b, st, 0.0, 0.0, 0, grid[kTabCell]);
    if (part == kPartAdvance || part == kPartFlip || part == kPartSeedDone) {
        const int g = (nslot_cap + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_solves, dim3(g > 0 ? g : 1), dim3(kBlock), 0, st, b, part);
        return hipGetLastError();
    }
    if (pde == kPdeAd) return launch_part_ad(N, b, st, part, grid);
    return pde == kPdeBn ? launch_part_pde<BnPDE>(N, b, st, part, grid, variant, upd)
                         : launch_part_pde<EulerPDE>(N, b, st, part, grid, variant, upd);
}

hipError_t launch_dt(const StepBufs &b, hipStream_t st, double time, double tend, int use_host_time, int grid)
{
    grid = grid < 1 ? 1 : (grid > kDtGrid ? kDtGrid : grid);
    hipLaunchKernelGGL(k_dt, dim3(grid), dim3(kBlock), 0, st, b, time, tend, use_host_time);
    return hipGetLastError();
}

hipError_t launch_tables(const StepBufs &b, const int *prefix, int nslot, const int quanta[kNumTabs], int max_entries,
                         hipStream_t st)
{
    int g = (max_entries + kBlock - 1) / kBlock;
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    hipLaunchKernelGGL(k_tables, dim3(g), dim3(kBlock), 0, st, b, prefix, nslot,
                       make_int4(quanta[0], quanta[1], quanta[2], quanta[3]));
    return hipGetLastError();
}

hipError_t launch_seed_ic(int pde, int N, const StepBufs &b, const IcSpec &ic, int slot, int ncell, hipStream_t st)
{
    const dim3 grid((unsigned)(ncell < 65536 ? ncell : 65536)), block(64);
    if (pde == kPdeBn) {
        switch (N) {
        case 1: hipLaunchKernelGGL((k_seed_ic<BnPDE, 1>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 2: hipLaunchKernelGGL((k_seed_ic<BnPDE, 2>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 3: hipLaunchKernelGGL((k_seed_ic<BnPDE, 3>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 4: hipLaunchKernelGGL((k_seed_ic<BnPDE, 4>), grid, block, 0, st, b, ic, slot, ncell); break;
        default: return hipErrorInvalidValue;
        }
    } else if (pde == kPdeEuler) {
        switch (N) {
        case 1: hipLaunchKernelGGL((k_seed_ic<EulerPDE, 1>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 2: hipLaunchKernelGGL((k_seed_ic<EulerPDE, 2>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 3: hipLaunchKernelGGL((k_seed_ic<EulerPDE, 3>), grid, block, 0, st, b, ic, slot, ncell); break;
        case 4: hipLaunchKernelGGL((k_seed_ic<EulerPDE, 4>), grid, block, 0, st, b, ic, slot, ncell); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_init(int pde, int N, const StepBufs &b, const IcSpec &ic, int slot, int ncell, int parity,
                       hipStream_t st)
{
    const long pts = (long)ncell * (N + 1) * (N + 1);
    const dim3 grid((unsigned)((pts + kBlock - 1) / kBlock)), block(kBlock);
    if (pde == kPdeAd) {
        switch (N) {
        case 1: hipLaunchKernelGGL((k_init<AdPDE, 1>), grid, block, 0, st, b, ic, slot, parity); break;
        case 2: hipLaunchKernelGGL((k_init<AdPDE, 2>), grid, block, 0, st, b, ic, slot, parity); break;
        case 3: hipLaunchKernelGGL((k_init<AdPDE, 3>), grid, block, 0, st, b, ic, slot, parity); break;
        case 4: hipLaunchKernelGGL((k_init<AdPDE, 4>), grid, block, 0, st, b, ic, slot, parity); break;
        default: return hipErrorInvalidValue;
        }
    } else if (pde == kPdeBn) {
        switch (N) {
        case 1: hipLaunchKernelGGL((k_init<BnPDE, 1>), grid, block, 0, st, b, ic, slot, parity); break;
        case 2: hipLaunchKernelGGL((k_init<BnPDE, 2>), grid, block, 0, st, b, ic, slot, parity); break;
        case 3: hipLaunchKernelGGL((k_init<BnPDE, 3>), grid, block, 0, st, b, ic, slot, parity); break;
        case 4: hipLaunchKernelGGL((k_init<BnPDE, 4>), grid, block, 0, st, b, ic, slot, parity); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (N) {
        case 1: hipLaunchKernelGGL((k_init<EulerPDE, 1>), grid, block, 0, st, b, ic, slot, parity); break;
        case 2: hipLaunchKernelGGL((k_init<EulerPDE, 2>), grid, block, 0, st, b, ic, slot, parity); break;
        case 3: hipLaunchKernelGGL((k_init<EulerPDE, 3>), grid, block, 0, st, b, ic, slot, parity); break;
        case 4: hipLaunchKernelGGL((k_init<EulerPDE, 4>), grid, block, 0, st, b, ic, slot, parity); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_test_predictor(int N, int variant, const MeshArgs &m, const StepBufs &b, const TestBufs &t,
                                 double dt, hipStream_t st)
{
    return m.pde == kPdeBn ? launch_test_pde<BnPDE>(N, m, b, t, dt, st, variant)
                           : launch_test_pde<EulerPDE>(N, m, b, t, dt, st, variant);
}

}  // namespace aderdg
