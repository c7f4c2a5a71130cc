// percell.hip -- per-cell forms of the reference's element routines for callers that keep
// the reference's loops over cells: the symbol-compatible Fortran shim
// (ader-dg_amd/fortran/aderdg_percell.f90) calls them once per cell from the reference's
// unedited main program.  Each kernel reads and writes the routine's own arguments in the
// reference layout, one cell after another, and takes every sum in the reference's order, so
// the build without contraction is bit-identical to the reference routines:
//
//   k_volume_cells   volume_integral   Euler/volume_integral.f90:8-36 (+ lShi,
//                                      Baer-Nunziato/volume_integral.f90)
//   k_surface_cells  surface_integral  Euler/surface_integral.f90:8-39
//   k_update_cells   element_update    Euler/element_update.f90:8-30
//   k_dmp_cells      DMP               Euler/limiter.f90:179-245 (PDEAssurePositivity of the PDE)
//
// One thread per (cell, node) for the first three, one per cell for the DMP: these entries
// carry a host round trip per call and are not the stepping path, which fuses the same
// arithmetic into k_predictor* (volume integral) and k_update (the rest).
#include "percell.hpp"
#include "pde_device.hpp"

namespace aderdg {
namespace {

constexpr int kCellBlock = 64;

template <int V, int N, bool SRC>
__global__ __launch_bounds__(kCellBlock) void k_volume_cells(MeshArgs m, const Operators *op, int ncells, CellBufs c)
{
    constexpr int n = N + 1, NN = n * n;
    const int t = blockIdx.x * kCellBlock + threadIdx.x;
    if (t >= ncells * NN) return;
    const int cell = t / NN, tt = t - cell * NN, i = tt % n, j = tt / n;
    const double *F = c.fhi + (size_t)cell * 2 * V * NN;  // lFhi(V, 2, n, n)
    double *d = c.lduh + (size_t)cell * V * NN;
#pragma unroll
    for (int v = 0; v < V; v++) {
        double l = 0.0;  // lduh = 0 (:18)
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < n; k++) acc = acc + F[v + V * (0 + 2 * (k + n * j))] * op->kxi[i + n * k];
        l = l + acc * op->w[j] / m.dx;  // x direction (:22-24)
        acc = 0.0;
#pragma unroll
        for (int k = 0; k < n; k++) acc = acc + F[v + V * (1 + 2 * (i + n * k))] * op->kxi[j + n * k];
        l = l + acc * op->w[i] / m.dy;  // y direction (:28-30)
        if (SRC) l = l + op->w[i] * op->w[j] * c.shi[(size_t)cell * V * NN + v + V * tt];
        d[v + V * tt] = l;
    }
}

template <int V, int N>
__global__ __launch_bounds__(kCellBlock) void k_surface_cells(MeshArgs m, const Operators *op, int ncells, CellBufs c)
{
    constexpr int n = N + 1, NN = n * n;
    const int t = blockIdx.x * kCellBlock + threadIdx.x;
    if (t >= ncells * NN) return;
    const int cell = t / NN, tt = t - cell * NN, i = tt % n, j = tt / n;
    const size_t f = (size_t)cell * V * n;  // lFl, lFr, lGb, lGt (V, n)
    double *d = c.lduh + (size_t)cell * V * NN;
    const double sx = op->w[j] / m.dx, sy = op->w[i] / m.dy;
#pragma unroll
    for (int v = 0; v < V; v++) {
        double l = d[v + V * tt];
        l = l - sx * (c.fr[f + v + V * j] * op->frc[i] - c.fl[f + v + V * j] * op->flc[i]);  // x faces (:21-26)
        l = l - sy * (c.gt[f + v + V * i] * op->frc[j] - c.gb[f + v + V * i] * op->flc[j]);  // y faces (:30-35)
        d[v + V * tt] = l;
    }
}

template <int V, int N>
__global__ __launch_bounds__(kCellBlock) void k_update_cells(const Operators *op, int ncells, CellBufs c, double dt)
{
    constexpr int n = N + 1, NN = n * n;
    const int t = blockIdx.x * kCellBlock + threadIdx.x;
    if (t >= ncells * NN) return;
    const int cell = t / NN, tt = t - cell * NN, i = tt % n, j = tt / n;
    double *d = c.lduh + (size_t)cell * V * NN, *u = c.luh + (size_t)cell * V * NN;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const double l = d[v + V * tt] / (op->w[i] * op->w[j]);  // :20-24
        d[v + V * tt] = l;
        u[v + V * tt] = u[v + V * tt] + dt * l;  // :28
    }
}

template <class P, int N>
__global__ __launch_bounds__(kCellBlock) void k_dmp_cells(MeshArgs m, const Operators *op, int ncells, CellBufs c,
                                                         double argmax)
{
    constexpr int V = P::V, n = N + 1, NN = n * n, s = 2 * N + 1, SS = s * s;
    const int cell = blockIdx.x * kCellBlock + threadIdx.x;
    if (cell >= ncells) return;
    const double *u = c.luh + (size_t)cell * V * NN;  // arguh (V, n, n)
    double *lim = c.work + (size_t)cell * V * (SS + NN), *lob = lim + V * SS;
    // get_subcell_data (Euler/limiter.f90:12-37) and get_lobatto_data (:43-68): first the
    // y contraction of a column a for row q, then the x contraction over a
    for (int q = 0; q < s; q++)
        for (int v = 0; v < V; v++) {
            double ly[n];
#pragma unroll
            for (int a = 0; a < n; a++) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < n; b++) acc = acc + u[v + V * (a + n * b)] * op->u2l[q + s * b];
                ly[a] = acc;
            }
            for (int p = 0; p < s; p++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + ly[a] * op->u2l[p + s * a];
                lim[v + V * (p + s * q)] = acc;
            }
        }
    for (int q = 0; q < n; q++)
        for (int v = 0; v < V; v++) {
            double ly[n];
#pragma unroll
            for (int a = 0; a < n; a++) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < n; b++) acc = acc + u[v + V * (a + n * b)] * op->u2b[q + n * b];
                ly[a] = acc;
            }
#pragma unroll
            for (int p = 0; p < n; p++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < n; a++) acc = acc + ly[a] * op->u2b[p + n * a];
                lob[v + V * (p + n * q)] = acc;
            }
        }
    int ok = 1;
    // the cell's range against the bounds, :196-222: min(minval(uh), minval(lim), minval(lob))
    for (int v = 0; v < V && ok; v++) {
        const double nan = __builtin_nan("");
        double u0 = nan, u1 = nan, l0 = nan, l1 = nan, b0 = nan, b1 = nan;
        for (int k = 0; k < NN; k++) {
            u0 = fold_min(u0, u[v + V * k]);
            u1 = fold_max(u1, u[v + V * k]);
            b0 = fold_min(b0, lob[v + V * k]);
            b1 = fold_max(b1, lob[v + V * k]);
        }
        for (int k = 0; k < SS; k++) {
            l0 = fold_min(l0, lim[v + V * k]);
            l1 = fold_max(l1, lim[v + V * k]);
        }
        const double mn = f_min(f_min(u0, l0), b0), mx = f_max(f_max(u1, l1), b1);
        const double lo = c.lmin[(size_t)cell * V + v], hi = c.lmax[(size_t)cell * V + v];
        const double ldiff = f_max(f_max(1.0e-4, 1.0e-3 * (hi - lo)), argmax);
        if (mn < lo - ldiff || mx > hi + ldiff) ok = 0;
    }
    // physical admissibility of every subcell value, :226-241
    for (int k = 0; k < SS && ok; k++) {
        double Q[V];
#pragma unroll
        for (int v = 0; v < V; v++) Q[v] = lim[v + V * k];
        if (P::inadmissible(m, Q)) ok = 0;
    }
    c.result[cell] = ok;
}

template <class P, int N>
hipError_t launch_impl(int kind, const MeshArgs &m, const Operators *op, int ncells, const CellBufs &c, double dt,
                       double argmax, hipStream_t st)
{
    constexpr int V = P::V, NN = (N + 1) * (N + 1);
    const int gnode = (ncells * NN + kCellBlock - 1) / kCellBlock, gcell = (ncells + kCellBlock - 1) / kCellBlock;
    switch (kind) {
    case kCellVolume:
        if (m.pde == kPdeBn)
            hipLaunchKernelGGL((k_volume_cells<V, N, true>), dim3(gnode), dim3(kCellBlock), 0, st, m, op, ncells, c);
        else
            hipLaunchKernelGGL((k_volume_cells<V, N, false>), dim3(gnode), dim3(kCellBlock), 0, st, m, op, ncells, c);
        break;
    case kCellSurface:
        hipLaunchKernelGGL((k_surface_cells<V, N>), dim3(gnode), dim3(kCellBlock), 0, st, m, op, ncells, c);
        break;
    case kCellUpdate:
        hipLaunchKernelGGL((k_update_cells<V, N>), dim3(gnode), dim3(kCellBlock), 0, st, op, ncells, c, dt);
        break;
    case kCellDmp:
        hipLaunchKernelGGL((k_dmp_cells<P, N>), dim3(gcell), dim3(kCellBlock), 0, st, m, op, ncells, c, argmax);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_pde(int kind, const MeshArgs &m, const Operators *op, int N, int ncells, const CellBufs &c,
                      double dt, double argmax, hipStream_t st)
{
    switch (N) {
    case 1: return launch_impl<P, 1>(kind, m, op, ncells, c, dt, argmax, st);
    case 2: return launch_impl<P, 2>(kind, m, op, ncells, c, dt, argmax, st);
    case 3: return launch_impl<P, 3>(kind, m, op, ncells, c, dt, argmax, st);
    case 4: return launch_impl<P, 4>(kind, m, op, ncells, c, dt, argmax, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_cells(int kind, const MeshArgs &m, const Operators *op, int N, int ncells, const CellBufs &c,
                        double dt, double argmax, hipStream_t st)
{
    if (ncells < 1) return hipSuccess;
    if (m.pde == kPdeBn) return launch_pde<BnPDE>(kind, m, op, N, ncells, c, dt, argmax, st);
    if (m.pde == kPdeEuler) return launch_pde<EulerPDE>(kind, m, op, N, ncells, c, dt, argmax, st);
    return hipErrorInvalidValue;
}

}  // namespace aderdg
